"""Pins of the CPU oracle (oracle/qcs_oracle.cpp) without the reference binary.

The reference ships no tests, fixtures or golden vectors, and cannot be built here (CUDA,
MPI, NCCL, atlc), so the oracle is pinned by:
  * the textbook model (tests/textbook.py): every gate, random targets, positive and
    negative controls, random input states and random qubit mappings;
  * closed forms of the reference's example circuits (GHZ, H on all, H on q0);
  * invariants: norm preservation, mapping invariance (bit-identical), U U^dagger = I;
  * exact arithmetic for the sampling prefix (dyadic probabilities).
"""
import math

import numpy as np
import pytest

import pyqcs
from pyqcs import circuits
import textbook
from conftest import random_state, fidelity

GATES_1 = ["h", "x", "y", "z", "s", "sdg", "t", "tdg", "sx", "sxdg", "id", "rx", "ry", "rz",
           "u1", "p", "u2", "u3", "u", "u4", "r"]
GATES_2 = ["swap", "iswap", "rzz", "rxx", "ryy", "rzx", "xx_plus_yy", "xx_minus_yy", "dcx", "ecr"]
GATES_34 = ["rccx", "rcccx"]
ALL = GATES_1 + GATES_2 + GATES_34 + ["global_phase"]


def _params(name, rng):
    k = 1 if name == "global_phase" else pyqcs.GATE_PARAMS[name]
    return tuple(rng.uniform(-math.pi, math.pi, size=k))


@pytest.mark.parametrize("name", ALL)
def test_textbook_matrices_unitary(name):
    rng = np.random.default_rng(1)
    m = textbook.matrix(name, _params(name, rng))
    assert np.allclose(m @ m.conj().T, np.eye(m.shape[0]), atol=1e-14)


def test_textbook_known_entries():
    # a few entries fixed by the published definitions, so a transposed or conjugated
    # textbook matrix fails here as well
    assert textbook.matrix("rccx")[3, 7] == -1j and textbook.matrix("rccx")[7, 3] == 1j
    assert textbook.matrix("rcccx")[7, 15] == 1 and textbook.matrix("rcccx")[15, 7] == -1
    assert np.isclose(textbook.matrix("ecr")[0, 3], 1j / math.sqrt(2))
    # DCX = cx(q0->q1) then cx(q1->q0): q0=1,q1=0 -> q0=1,q1=1 -> q0=0,q1=1 (index 2)
    v = np.zeros(4)
    v[1] = 1
    assert np.argmax(abs(textbook.matrix("dcx") @ v)) == 2
    # RZX(pi) = -i Z_q0 X_q1: |00> -> -i |10> (q1 flipped)
    v = np.zeros(4)
    v[0] = 1
    out = textbook.matrix("rzx", [math.pi]) @ v
    assert np.isclose(out[2], -1j)
    # U3(theta, 0, 0) = RY(theta); U2(phi, lambda) = U3(pi/2, phi, lambda)
    assert np.allclose(textbook.matrix("u3", [0.3, 0, 0]), textbook.matrix("ry", [0.3]))


@pytest.mark.parametrize("name", ALL)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_gate_matches_textbook(oracle, name, seed):
    n = 6
    rng = np.random.default_rng(100 + seed + 7 * ALL.index(name))
    t = 0 if name == "global_phase" else pyqcs.GATE_TARGETS[name]
    qs = rng.permutation(n)
    targets = [int(q) for q in qs[:t]]
    nc = rng.integers(0, 3)
    ctl = [int(q) for q in qs[t:t + nc]]
    pos = [c for c in ctl if rng.random() < 0.5]
    neg = [c for c in ctl if c not in pos]
    params = _params(name, rng)
    mapping = [int(q) for q in rng.permutation(n)] if seed == 2 else None
    psi = random_state(n, seed)
    with pyqcs.Simulator(oracle, n, mapping=mapping) as sim:
        sim.set_statevector(psi)
        sim.gate(name, targets, params, pos, neg)
        got = sim.statevector()
    want = textbook.apply(psi, n, name, targets, params, pos, neg)
    assert np.max(np.abs(got - want)) < 1e-14


def test_oracle_random_circuit_matches_textbook(oracle):
    n = 7
    circ = circuits.random_all_gates(n, 300, seed=11)
    with pyqcs.Simulator(oracle, n) as sim:
        sim.run(circ)
        got = sim.statevector()
    want = textbook.run(n, circ)
    assert np.max(np.abs(got - want)) < 1e-13
    assert fidelity(got, want) > 1 - 1e-12


def test_mapping_invariance_bit_identical(oracle):
    n = 8
    circ = circuits.random_all_gates(n, 200, seed=5)
    outs = []
    for mapping in [None, list(range(n))[::-1], [3, 7, 1, 0, 6, 2, 5, 4]]:
        with pyqcs.Simulator(oracle, n, mapping=mapping) as sim:
            sim.run(circ)
            outs.append(sim.statevector())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_closed_forms(oracle):
    n = 9
    with pyqcs.Simulator(oracle, n) as sim:  # examples/standalone/ghz_circuit.c:30-41
        sim.run(circuits.ghz_u3_cz(n, 0, 0))
        psi = sim.statevector()
    want = np.zeros(1 << n, complex)
    want[0] = want[-1] = 1 / math.sqrt(2)
    assert np.max(np.abs(psi - want)) < 1e-15
    with pyqcs.Simulator(oracle, n) as sim:  # hadamard-whole.c:29-35
        for q in range(n):
            sim.gate("h", [q])
        assert np.allclose(sim.statevector(), 2 ** (-n / 2), atol=1e-15)
    with pyqcs.Simulator(oracle, n) as sim:  # hadamard-single.c:29-33
        sim.gate("h", [0])
        psi = sim.statevector()
        assert np.isclose(psi[0], 1 / math.sqrt(2)) and np.isclose(psi[1], 1 / math.sqrt(2))
        assert np.count_nonzero(psi) == 2


def test_qft_closed_form(oracle):
    # QFT of |x> has amplitudes exp(2 pi i x k / N) / sqrt(N) (with the final swaps)
    n = 6
    N = 1 << n
    x = 13
    psi = np.zeros(N, complex)
    psi[x] = 1
    with pyqcs.Simulator(oracle, n) as sim:
        sim.set_statevector(psi)
        sim.run(circuits.qft(n))
        got = sim.statevector()
    k = np.arange(N)
    want = np.exp(2j * np.pi * x * k / N) / math.sqrt(N)
    assert np.max(np.abs(got - want)) < 1e-14


def test_norm_preserved_and_canonical_sum(oracle):
    n = 12
    with pyqcs.Simulator(oracle, n) as sim:
        sim.run(circuits.brick(n, 6, seed=3, mcx=2))
        assert abs(sim.norm() - 1) < 1e-13
        assert abs(sim.probability_total() - 1) < 1e-13


def canonical_sum_numpy(x):
    """The canonical order of include/qcs_ext.h, written with numpy slicing."""
    x = np.asarray(x, dtype=np.float64)
    while True:
        pad = (-x.size) % 4096
        t = np.concatenate([x, np.zeros(pad)]).reshape(-1, 4096)
        while t.shape[1] > 1:
            h = t.shape[1] // 2
            t = t[:, :h] + t[:, h:]
        x = t[:, 0]
        if x.size == 1:
            return float(x[0])


def test_canonical_sum_order(oracle):
    # p = 1 at index 0 of each of two tiles and 2^-60 elsewhere: a left-to-right sum
    # gives exactly 2.0; the tree keeps part of the small terms.
    n = 13
    N = 1 << n
    amp = np.full(N, 2.0 ** -30, complex)
    amp[0] = 1.0
    amp[4096] = 1.0
    p = np.abs(amp) ** 2
    with pyqcs.Simulator(oracle, n) as sim:
        sim.set_statevector(amp)
        got = sim.norm()
    assert got == canonical_sum_numpy(p)
    assert got != 2.0
    assert abs(got - math.fsum(p)) < 1e-15
    rng = np.random.default_rng(4)
    psi = random_state(14, 4) * rng.uniform(0.5, 2)
    with pyqcs.Simulator(oracle, 14) as sim:
        sim.set_statevector(psi)
        assert sim.norm() == canonical_sum_numpy(psi.real * psi.real + psi.imag * psi.imag) or \
            abs(sim.norm() - canonical_sum_numpy(np.abs(psi) ** 2)) < 1e-15


def test_sampling_exact_inverse_cdf(oracle):
    # amplitudes sqrt(w) * 2^-m with w in {0,1,4,9}: every p and every prefix sum is an
    # exact dyadic number, so the right-nested prefix equals the plain cumulative sum and
    # the outcome must be searchsorted(cumsum, r, 'left')
    n = 10
    N = 1 << n
    rng = np.random.default_rng(7)
    w = rng.choice([0, 1, 4, 9], size=N).astype(float)
    w[-3:] = 0
    m = math.ceil(math.log(w.sum(), 4))
    amp = (np.sqrt(w) * 2.0 ** -m).astype(complex)
    p = amp.real * amp.real
    assert np.array_equal(p, w * 4.0 ** -m)
    shots = 4000
    with pyqcs.Simulator(oracle, n) as sim:
        sim.set_statevector(amp)
        out = sim.sample(shots, seed=99)
        total = sim.probability_total()
    assert total == p.sum()
    u = np.array(circuits.uniform_stream(99, shots))
    want = np.searchsorted(np.cumsum(p), u * total, side="left")
    assert np.array_equal(out, want)
    assert np.all(p[out] > 0)


def test_sampling_distribution(oracle):
    n = 6
    psi = random_state(n, 3)
    shots = 60000
    with pyqcs.Simulator(oracle, n) as sim:
        sim.set_statevector(psi)
        out = sim.sample(shots, seed=5)
    counts = np.bincount(out.astype(np.int64), minlength=1 << n)
    p = np.abs(psi) ** 2
    chi2 = np.sum((counts - shots * p) ** 2 / (shots * p))
    assert chi2 < 64 + 6 * math.sqrt(2 * 64)


def test_sampling_reports_logical_index(oracle):
    n = 5
    mapping = [4, 3, 2, 1, 0]
    with pyqcs.Simulator(oracle, n, mapping=mapping) as sim:
        sim.gate("x", [1])  # logical basis state 0b00010
        out = sim.sample(10, seed=1)
    assert np.all(out == 2)


def test_measure_ghz_and_lazy_controls(oracle):
    n = 8
    for seed in range(6):
        with pyqcs.Simulator(oracle, n, num_clbits=n, seed=seed) as sim:
            sim.run(circuits.ghz_u3_cz(n, 0, 0))
            bits = [sim.measure_to_clbit(q, q) for q in range(n)]
            assert len(set(bits)) == 1  # all 0 or all 1
            s = sim.clbits_string()
            assert s == "".join(str(b) for b in reversed(bits))
            psi = sim.statevector()
            idx = 0 if bits[0] == 0 else (1 << n) - 1
            # collapsed but not renormalised (src/qcs.cu:1106-1123)
            assert np.isclose(abs(psi[idx]), 1 / math.sqrt(2))
            assert np.count_nonzero(psi) == 1


class MT19937_64:
    """Matsumoto & Nishimura's 64-bit Mersenne Twister, as specified for std::mt19937_64."""

    def __init__(self, seed):
        M = (1 << 64) - 1
        self.mt = [seed & M]
        for i in range(1, 312):
            prev = self.mt[-1]
            self.mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & M)
        self.i = 312

    def __call__(self):
        M = (1 << 64) - 1
        if self.i >= 312:
            mt = self.mt
            for k in range(312):
                y = (mt[k] & 0xFFFFFFFF80000000) | (mt[(k + 1) % 312] & 0x7FFFFFFF)
                v = mt[(k + 156) % 312] ^ (y >> 1)
                if y & 1:
                    v ^= 0xB5026F5AA96619E9
                mt[k] = v
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & M


def test_mt19937_64_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937_64
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


def test_measure_rule_exact(oracle):
    # src/qcs.cu:2615-2617: r = uniform_real_distribution<double>(0, p0+p1)(engine),
    # result = p0 < r; the libstdc++ canonical draw is double(x) / 2^64 for one 64-bit word
    for seed in range(60):
        theta = 0.05 + 3.0 * seed / 60
        with pyqcs.Simulator(oracle, 1, seed=seed) as sim:
            sim.gate("ry", [0], [theta])
            a = sim.statevector()
            got = sim.measure(0)
        assert a.imag[0] == 0 and a.imag[1] == 0
        p0, p1 = a.real[0] * a.real[0], a.real[1] * a.real[1]
        r = (p0 + p1) * (float(MT19937_64(seed)()) / 2.0 ** 64)
        assert got == int(p0 < r), seed
    # tie rule: a zero-probability |1> branch is never chosen, p0 = 1 gives 0
    for seed in range(10):
        with pyqcs.Simulator(oracle, 2, seed=seed) as sim:
            assert sim.measure(1) == 0


def test_lazy_measurement_semantics(oracle):
    n = 4
    with pyqcs.Simulator(oracle, n, seed=3) as sim:
        sim.gate("h", [0])
        sim.gate("h", [1])
        m0 = sim.measure(0)
        # cached: no new draw, same value
        assert sim.measure(0) == m0
        before = sim.statevector()
        # a gate whose explicit control contradicts the measured value is skipped
        if m0 == 1:
            sim.gate("x", [2], negctrls=[0])
        else:
            sim.gate("x", [2], ctrls=[0])
        assert np.array_equal(sim.statevector(), before)
        # a gate on another qubit acts only on the observed branch (implicit control)
        sim.gate("x", [3])
        psi = sim.statevector()
        for i in range(1 << n):
            if ((i >> 0) & 1) != m0:
                assert psi[i] == 0
        # targeting the measured qubit releases it from the measured list
        sim.gate("h", [0])
        psi = sim.statevector()
        assert np.count_nonzero(np.abs(psi) > 1e-15) == 4


def test_reset(oracle):
    n = 3
    for seed in range(5):
        with pyqcs.Simulator(oracle, n, seed=seed) as sim:
            sim.gate("h", [1])
            sim.reset(1)
            psi = sim.statevector()
            assert np.all(np.abs(psi[[i for i in range(8) if (i >> 1) & 1]]) < 1e-15)


def test_renormalize(oracle):
    n = 5
    with pyqcs.Simulator(oracle, n, seed=2) as sim:
        sim.set_renormalize(True)
        for q in range(n):
            sim.gate("h", [q])
        for q in range(n):
            sim.measure(q)
        assert abs(np.linalg.norm(sim.statevector()) - 1) < 1e-14


@pytest.mark.parametrize("mapping", [None, [5, 2, 7, 0, 3, 6, 1, 4]])
def test_state_preps_physical_index(oracle, mapping):
    # the reference fills the device array by physical index (src/qcs.cu:71-82, 2025-2095):
    # sequential puts value P at physical index P, so logical index L reads phys(L) =
    # sum_q bit(L, q) << l2p[q]; flat is all ones; entangled (one rank) sets physical index
    # 0 to 1 and clears nothing
    n = 8
    l2p = list(range(n)) if mapping is None else [mapping.index(q) for q in range(n)]
    L = np.arange(1 << n)
    phys = np.zeros(1 << n, dtype=np.int64)
    for q in range(n):
        phys |= ((L >> q) & 1) << l2p[q]
    with pyqcs.Simulator(oracle, n, mapping=mapping) as sim:
        sim.call("set_sequential_state")
        assert np.array_equal(sim.statevector(), phys.astype(np.complex128))
        sim.call("set_entangled_state")
        want = phys.astype(np.complex128)
        want[0] = 1.0
        assert np.array_equal(sim.statevector(), want)
        sim.call("set_flat_state")
        assert np.array_equal(sim.statevector(), np.ones(1 << n, dtype=np.complex128))
        sim.call("set_zero_state")
        sim.gate("h", [3])
        sim.call("set_entangled_state")
        want = np.zeros(1 << n, dtype=np.complex128)
        want[0] = 1.0
        want[1 << 3] = 1 / math.sqrt(2)
        assert np.max(np.abs(sim.statevector() - want)) <= 2e-16


def test_renormalised_collapse_closed_form(oracle):
    # after measure(q) = b with renormalisation: psi restricted to bit q == b over the
    # square root of that branch's probability (the p0/p1 of the measurement)
    n = 7
    psi = random_state(n, 31)
    with pyqcs.Simulator(oracle, n, seed=4, mapping=[6, 0, 5, 1, 4, 2, 3]) as sim:
        sim.set_renormalize(True)
        sim.set_statevector(psi)
        want = psi.copy()
        for q in (2, 0, 6):
            b = sim.measure(q)
            keep = ((np.arange(1 << n) >> q) & 1) == b
            want = np.where(keep, want, 0)
            want = want / math.sqrt(float(np.sum(np.abs(want) ** 2)))
            assert np.max(np.abs(sim.statevector() - want)) <= 1e-15


def test_local_window_swap_in(oracle):
    # src/qcs.cu:2216-2226: a target on global physical qubit a_i is swapped with local
    # slot n_local-1-i; maps updated as at src/qcs.cu:2306-2328
    n = 6
    circ = circuits.random_all_gates(n, 80, seed=21)
    with pyqcs.Simulator(oracle, n) as ref:
        ref.run(circ)
        want = ref.statevector()
    with pyqcs.Simulator(oracle, n) as sim:
        sim.set_local_qubits(4)
        sim.gate("h", [5])
        assert sim.mapping() == [0, 1, 2, 5, 4, 3]
        sim.gate("h", [5])  # now local: no further swap
        assert sim.mapping() == [0, 1, 2, 5, 4, 3]
        sim.gate("h", [5])
        sim.gate("h", [5])
        sim.call("set_zero_state")
        sim.run(circ)
        got = sim.statevector()
    assert np.max(np.abs(got - want)) < 1e-14


def test_permute_keeps_logical_state(oracle):
    n = 7
    psi = random_state(n, 8)
    with pyqcs.Simulator(oracle, n) as sim:
        sim.set_statevector(psi)
        sim.permute([6, 0, 5, 1, 4, 2, 3])
        assert sim.mapping() == [6, 0, 5, 1, 4, 2, 3]
        assert np.array_equal(sim.statevector(), psi)


def test_save_statevector_format(oracle, tmp_path):
    n = 7
    mapping = [2, 0, 1, 6, 5, 3, 4]
    circ = circuits.random_all_gates(n, 50, seed=9)
    with pyqcs.Simulator(oracle, n, mapping=mapping) as sim:
        sim.run(circ)
        path = tmp_path / "sv.bin"
        sim.save(path)
        psi = sim.statevector()
    raw = np.fromfile(path, dtype="<f8")
    assert raw.size == 2 << n
    assert np.array_equal(raw[0::2] + 1j * raw[1::2], psi)
    assert np.max(np.abs(psi - textbook.run(n, circ))) < 1e-13


def test_error_convention(oracle):
    lib = oracle.lib
    import ctypes
    assert lib.qcs_simulator_create(None) == 0
    assert "NULL" in oracle.last_error
    with pyqcs.Simulator(oracle, 3, num_clbits=2) as sim:
        h = sim.h
        assert lib.qcs_simulator_get_num_qubits(h, None) == 0
        t = (pyqcs.bit_num_t * 1)(0)
        assert lib.qcs_simulator_gate_h(h, t, -1, None, 0, None, 0) == 0
        assert lib.qcs_simulator_gate_h(h, None, 1, None, 0, None, 0) == 0
        assert lib.qcs_simulator_gate_swap(h, t, 1, None, 0, None, 0) == 0  # wrong arity
        bad = (pyqcs.bit_num_t * 1)(3)
        assert lib.qcs_simulator_gate_h(h, bad, 1, None, 0, None, 0) == 0  # out of range
        assert lib.qcs_simulator_gate_x(h, t, 1, None, 0, t, 1) == 0  # duplicate operand
        r = pyqcs.bit_t()
        assert lib.qcs_simulator_measure_to_clbit(h, 0, 5, ctypes.byref(r)) == 0
        assert lib.qcs_simulator_allocate_memory(h) == 0  # twice
        assert "already" in oracle.last_error or oracle.last_error == ""
        assert lib.qcs_simulator_save_statevector(h, b"/nonexistent-dir/x.bin") == 0
    with pyqcs.Simulator(oracle, 3, allocate=False) as sim:
        for m in ([0, 1], [0, 1, 3], [0, 1, 1]):
            a = (pyqcs.bit_num_t * len(m))(*m)
            assert sim.lib.qcs_simulator_set_mapping(sim.h, a, len(m)) == 0


def test_events(oracle):
    with pyqcs.Simulator(oracle, 4) as sim:
        a, b = sim.event_create(), sim.event_create()
        sim.event_record(a)
        sim.gate("h", [0])
        sim.event_record(b)
        assert sim.elapsed(a, b) >= 0


def test_dispose_then_destroy_is_safe(oracle):
    sim = pyqcs.Simulator(oracle, 3)
    assert sim.lib.qcs_simulator_dispose(sim.h) == 1
    assert sim.lib.qcs_simulator_gate_h(sim.h, (pyqcs.bit_num_t * 1)(0), 1, None, 0, None, 0) == 0
    sim.close()
