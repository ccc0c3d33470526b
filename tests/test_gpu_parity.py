"""GPU parity: libqcs.so (HIP kernels) against the CPU oracle through the same C ABI.

Bars (written in each test): amplitudes max|d| <= 1e-13 and fidelity >= 1 - 1e-12 in FP64;
probability totals, reductions and sampled outcomes bit-exact when both sides hold the
same state bits; measurement outcomes equal for the same seed.
"""
import math
import sys

import numpy as np
import pytest

import pyqcs
from pyqcs import circuits
import textbook
from conftest import random_state, fidelity

pytestmark = pytest.mark.gpu

AMP_TOL = 1e-13
FID_TOL = 1e-12

ALL = sorted(pyqcs.GATE_TARGETS)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_both(gpu, oracle, n, circ, mapping=None, psi=None, seed=None):
    out = []
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, n, mapping=mapping, seed=seed) as sim:
            if psi is not None:
                sim.set_statevector(psi)
            sim.run(circ)
            out.append(sim.statevector())
    return out


@pytest.mark.parametrize("name", ALL)
def test_each_gate(gpu, oracle, name):
    n = 11
    rng = np.random.default_rng(ALL.index(name))
    for trial in range(4):
        t = pyqcs.GATE_TARGETS[name]
        qs = [int(q) for q in rng.permutation(n)]
        targets = qs[:t]
        nc = int(rng.integers(0, 4))
        ctl = qs[t:t + nc]
        pos = [c for c in ctl if rng.random() < 0.5]
        neg = [c for c in ctl if c not in pos]
        k = 1 if name == "global_phase" else pyqcs.GATE_PARAMS[name]
        params = tuple(rng.uniform(-math.pi, math.pi, size=k))
        mapping = [int(q) for q in rng.permutation(n)] if trial == 3 else None
        psi = random_state(n, trial)
        say(f"gate {name} trial {trial} targets={targets} pos={pos} neg={neg}")
        g, o = run_both(gpu, oracle, n, [(name, targets, params, pos, neg)], mapping=mapping, psi=psi)
        assert np.max(np.abs(g - o)) <= AMP_TOL
        want = textbook.apply(psi, n, name, targets, params, pos, neg)
        assert np.max(np.abs(g - want)) <= AMP_TOL


ONE_TARGET_DENSE = ["h", "x", "y", "sx", "sxdg", "rx", "ry", "r", "u", "u2", "u3", "u4"]


def test_low_target_pair_kernel(gpu, oracle):
    # targets 0..5 with controls at 6 and above (positive and negative) run in the
    # cross-lane kernel; a measured target zeroes one branch first
    n = 14
    rng = np.random.default_rng(77)
    psi = random_state(n, 9)
    circ = []
    for i in range(120):
        name = ONE_TARGET_DENSE[i % len(ONE_TARGET_DENSE)]
        t = int(rng.integers(0, 6))
        ctl = [int(q) for q in rng.choice(np.arange(6, n), size=int(rng.integers(0, 3)), replace=False)]
        pos = [c for c in ctl if rng.random() < 0.5]
        neg = [c for c in ctl if c not in pos]
        params = tuple(rng.uniform(-math.pi, math.pi, size=pyqcs.GATE_PARAMS[name]))
        circ.append((name, (t,), params, tuple(pos), tuple(neg)))
    res = []
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, n, num_clbits=2, seed=5) as sim:
            sim.set_statevector(psi)
            sim.run(circ[:60])
            bits = [sim.measure_to_clbit(2, 0), sim.measure_to_clbit(9, 1)]
            sim.run(circ[60:])
            res.append((sim.statevector(), bits))
    assert res[0][1] == res[1][1]
    assert np.max(np.abs(res[0][0] - res[1][0])) <= AMP_TOL
    assert fidelity(res[0][0], res[1][0]) >= 1 - FID_TOL


def test_low_operand_phase_kernel(gpu, oracle):
    # phase gates whose operands include qubits 0..2 run in the 64-amplitude chunk kernel;
    # controls positive and negative, low and high
    n = 14
    rng = np.random.default_rng(78)
    psi = random_state(n, 10)
    circ = []
    for i in range(120):
        name = ["z", "s", "sdg", "t", "tdg", "p", "u1"][i % 7]
        t = int(rng.integers(0, 3))
        others = [q for q in range(n) if q != t]
        ctl = [int(q) for q in rng.choice(others, size=int(rng.integers(0, 3)), replace=False)]
        pos = [c for c in ctl if rng.random() < 0.5]
        neg = [c for c in ctl if c not in pos]
        params = tuple(rng.uniform(-math.pi, math.pi, size=pyqcs.GATE_PARAMS[name]))
        circ.append((name, (t,), params, tuple(pos), tuple(neg)))
    for i in range(20):  # CZ across low and high qubits, both orders
        a, b = int(rng.integers(0, 3)), int(rng.integers(3, n))
        circ.append(("z", (a,), (), (b,), ()) if i % 2 else ("z", (b,), (), (a,), ()))
    g, o = run_both(gpu, oracle, n, circ, psi=psi)
    assert np.max(np.abs(g - o)) <= AMP_TOL
    assert fidelity(g, o) >= 1 - FID_TOL


def test_low_operand_gate_kernel(gpu, oracle):
    # every gate with an operand on qubit 0, 1 or 2: 2-4-target gates (swap-like, dense
    # 4x4, one nonzero per row) and 1-target gates with a control there run in the
    # 64-amplitude chunk kernel, with up to two targets at qubit 6 or above as registers;
    # controls low and high, positive and negative; a measured target zeroes a branch
    n = 14
    rng = np.random.default_rng(79)
    psi = random_state(n, 11)
    circ = []
    for i, name in enumerate(ALL * 3):
        t = pyqcs.GATE_TARGETS[name]
        if t == 0:
            continue
        low = [int(q) for q in rng.permutation(3)]
        mid = [int(q) for q in rng.permutation(np.arange(3, 6))]
        high = [int(q) for q in rng.permutation(np.arange(6, n))]
        nhigh = min(t - 1, int(rng.integers(0, 3)))
        targets = low[:1] + (mid + low[1:])[: t - 1 - nhigh] + high[:nhigh]
        targets = [int(q) for q in rng.permutation(targets)]
        rest = [q for q in low + mid + high if q not in targets]
        ctl = [int(q) for q in rng.choice(rest, size=int(rng.integers(0, 3)), replace=False)]
        if t == 1 and i % 2:  # target high, control low
            targets = [high[-1]]
            ctl = [low[0]]
        pos = [c for c in ctl if rng.random() < 0.5]
        neg = [c for c in ctl if c not in pos]
        k = pyqcs.GATE_PARAMS[name]
        params = tuple(rng.uniform(-math.pi, math.pi, size=k))
        circ.append((name, tuple(targets), params, tuple(pos), tuple(neg)))
    say(f"low-operand circuit n={n} {len(circ)} gates")
    res = []
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, n, num_clbits=2, seed=6) as sim:
            sim.set_statevector(psi)
            sim.run(circ[: len(circ) // 2])
            bits = [sim.measure_to_clbit(1, 0), sim.measure_to_clbit(7, 1)]
            sim.run(circ[len(circ) // 2:])
            res.append((sim.statevector(), bits))
    assert res[0][1] == res[1][1]
    assert np.max(np.abs(res[0][0] - res[1][0])) <= AMP_TOL
    assert fidelity(res[0][0], res[1][0]) >= 1 - FID_TOL
    # the textbook model gate by gate on the same circuit without the measurements
    g, o = run_both(gpu, oracle, n, circ, psi=psi)
    want = psi
    for name, targets, params, pos, neg in circ:
        want = textbook.apply(want, n, name, list(targets), params, list(pos), list(neg))
    assert np.max(np.abs(g - want)) <= AMP_TOL
    assert np.max(np.abs(g - o)) <= AMP_TOL


@pytest.mark.parametrize("renorm", [False, True])
def test_collapse_states(gpu, oracle, renorm):
    # post-measurement states (not only their norm) against the oracle: qubits 0..2 take
    # the whole-line collapse kernel, the others the pair-wise one; plain and renormalised
    n = 13
    psi = random_state(n, 12)
    res = []
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, n, seed=21, mapping=[3, 0, 12, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11]) as sim:
            sim.set_renormalize(renorm)
            sim.set_statevector(psi)
            bits = [sim.measure(q) for q in (1, 0, 9, 2, 12)]
            res.append((sim.statevector(), bits))
    assert res[0][1] == res[1][1]
    assert np.max(np.abs(res[0][0] - res[1][0])) <= AMP_TOL
    norm = float(np.vdot(res[0][0], res[0][0]).real)
    if renorm:
        assert abs(norm - 1) <= 1e-12
    else:
        assert norm < 1 - 1e-3  # five measurements of a random state discard weight


def test_random_circuit_all_gates(gpu, oracle):
    n = 12
    circ = circuits.random_all_gates(n, 400, seed=17, max_ctrls=3)
    say("random circuit n=12 400 gates")
    g, o = run_both(gpu, oracle, n, circ, mapping=[5, 3, 11, 0, 1, 10, 2, 4, 9, 6, 8, 7])
    assert np.max(np.abs(g - o)) <= AMP_TOL
    assert fidelity(g, o) >= 1 - FID_TOL


@pytest.mark.parametrize("n", [22, 24])
def test_cached_size_all_gates(gpu, oracle, n):
    # 2^22 and 2^24 amplitudes (64 and 256 MiB): the gate kernels' cached-access
    # instantiations (pair, low-target pair, phase, low-operand phase, diagonal); low and
    # high targets and controls, on a random state
    circ = circuits.random_all_gates(n, 120, seed=60 + n, max_ctrls=3)
    circ += [("h", (t,), (), (), ()) for t in (0, 3, 5, 12, n - 1)]
    circ += [("z", (1,), (), (0,), ()), ("s", (2,), (), (n - 1,), ()), ("rz", (n - 2,), (0.4,), (), ())]
    psi = random_state(n, 3)
    say(f"cached-size circuit n={n}")
    g, o = run_both(gpu, oracle, n, circ, psi=psi)
    assert np.max(np.abs(g - o)) <= AMP_TOL
    assert fidelity(g, o) >= 1 - FID_TOL


def test_config_ghz_u3_cz_20q(gpu, oracle):
    # 20 qubits: H, CNOT chain, 20 layers of U3 + CZ matching; statevector and 1024 samples
    n = 20
    circ = circuits.ghz_u3_cz(n, 20, seed=2024)
    say("ghz_u3_cz n=20")
    sv, samples, totals = [], [], []
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, n) as sim:
            sim.run(circ)
            sv.append(sim.statevector())
            samples.append(sim.sample(1024, seed=77))
            totals.append(sim.probability_total())
    assert np.max(np.abs(sv[0] - sv[1])) <= AMP_TOL
    assert fidelity(sv[0], sv[1]) >= 1 - FID_TOL
    assert abs(totals[0] - 1) < 1e-12
    # outcomes from each side's own state: equal unless a uniform falls within rounding of
    # a prefix boundary (probability ~1e-13 per shot)
    assert np.array_equal(samples[0], samples[1])


def test_sampling_bit_exact_on_identical_state(gpu, oracle):
    for n in (7, 13, 19):
        psi = random_state(n, n)
        say(f"sampling identical state n={n}")
        res = []
        for lib in (gpu, oracle):
            with pyqcs.Simulator(lib, n) as sim:
                sim.set_statevector(psi)
                res.append((sim.probability_total(), sim.norm(), sim.sample(5000, seed=n), sim.expect_z()))
        assert res[0][0] == res[1][0]  # prefix total: bit-exact
        assert res[0][1] == res[1][1]  # canonical sum: bit-exact
        assert np.array_equal(res[0][2], res[1][2])  # outcomes: bit-exact
        assert np.array_equal(res[0][3], res[1][3])  # <Z_q>: canonical sums, bit-exact


def test_measurement_sequences(gpu, oracle):
    n = 10
    circ = circuits.random_all_gates(n, 60, seed=3)
    for seed in range(5):
        say(f"measurement sequence seed {seed}")
        outs = []
        for lib in (gpu, oracle):
            with pyqcs.Simulator(lib, n, num_clbits=n, seed=seed) as sim:
                sim.run(circ[:30])
                a = [sim.measure_to_clbit(q, q) for q in (0, 3, 7)]
                sim.run(circ[30:])  # gates after measurement see implicit controls
                b = [sim.measure(q) for q in range(n)]
                sim.reset(2)
                outs.append((a, b, sim.clbits_string(), sim.statevector()))
        assert outs[0][0] == outs[1][0] and outs[0][1] == outs[1][1] and outs[0][2] == outs[1][2]
        assert np.max(np.abs(outs[0][3] - outs[1][3])) <= AMP_TOL


def test_renormalize(gpu, oracle):
    n = 9
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, n, seed=4) as sim:
            sim.set_renormalize(True)
            sim.run(circuits.brick(n, 3, seed=1))
            for q in range(0, n, 2):
                sim.measure(q)
            assert abs(np.linalg.norm(sim.statevector()) - 1) < 1e-13


def test_local_window_and_permute(gpu, oracle):
    n = 10
    circ = circuits.random_all_gates(n, 120, seed=8)
    outs = []
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, n, seed=1) as sim:
            sim.set_local_qubits(6)
            sim.run(circ)
            m = sim.mapping()
            r = [sim.measure(q) for q in (9, 8)]
            sim.permute(list(range(n)))
            outs.append((m, r, sim.statevector(), sim.mapping()))
    assert outs[0][0] == outs[1][0]
    assert outs[0][1] == outs[1][1]
    assert outs[0][3] == list(range(n))
    assert np.max(np.abs(outs[0][2] - outs[1][2])) <= AMP_TOL


def test_save_statevector(gpu, oracle, tmp_path):
    n = 14
    mapping = list(np.random.default_rng(0).permutation(n))
    circ = circuits.brick(n, 4, seed=5, mcx=2)
    paths = []
    for tag, lib in (("g", gpu), ("o", oracle)):
        with pyqcs.Simulator(lib, n, mapping=mapping) as sim:
            sim.run(circ)
            p = tmp_path / f"{tag}.bin"
            sim.save(p)
            paths.append(p)
    a = pyqcs.load_statevector_file(paths[0], n)
    b = pyqcs.load_statevector_file(paths[1], n)
    assert np.max(np.abs(a - b)) <= AMP_TOL


@pytest.mark.parametrize("mapping", [None, [5, 2, 7, 0, 3, 6, 1, 4]])
def test_state_preps(gpu, oracle, mapping):
    # physical-index fills (src/qcs.cu:71-82, 2025-2095): under a mapping the logical
    # vector differs from the identity mapping's, so both sides must write physical indices
    n = 8
    for what in ("set_sequential_state", "set_flat_state", "set_entangled_state", "set_zero_state"):
        outs = []
        for lib in (gpu, oracle):
            with pyqcs.Simulator(lib, n, mapping=mapping) as sim:
                sim.gate("h", [3])
                sim.call(what)
                outs.append(sim.statevector())
        assert np.array_equal(outs[0], outs[1]), what
        if what == "set_sequential_state" and mapping is not None:
            assert not np.array_equal(outs[0], np.arange(1 << n, dtype=np.complex128))


def test_events_and_errors(gpu):
    with pyqcs.Simulator(gpu, 12) as sim:
        a, b = sim.event_create(), sim.event_create()
        sim.event_record(a)
        for q in range(12):
            sim.gate("h", [q])
        sim.event_record(b)
        assert sim.elapsed(a, b) > 0
        with pytest.raises(pyqcs.QcsError):
            sim.gate("h", [12])
        with pytest.raises(pyqcs.QcsError):
            sim.gate("x", [1], ctrls=[1])


def test_capacity_check(gpu):
    sim = pyqcs.Simulator(gpu, 39, allocate=False)
    with pytest.raises(pyqcs.QcsError):
        sim.allocate()  # 8 TiB of complex128 does not fit in HBM
    sim.close()


def test_full_size_30q_properties(gpu):
    # 30 qubits (16 GiB): GHZ probabilities, norm, U U^dagger round trip on a brick layer
    n = 30
    say("30-qubit properties")
    with pyqcs.Simulator(gpu, n) as sim:
        sim.run(circuits.ghz_u3_cz(n, 0, 0))
        out = sim.sample(4096, seed=1)
        assert set(np.unique(out).tolist()) <= {0, (1 << n) - 1}
        assert abs(sim.norm() - 1) < 1e-12
        layer = circuits.brick(n, 2, seed=9, mcx=2)
        sim.run(layer)
        assert abs(sim.norm() - 1) < 1e-12
        # inverse: reverse order, U3(t,p,l)^-1 = U3(-t,-l,-p); CZ and MCX are self-inverse
        inv = []
        for name, t, prm, c, ng in reversed(layer):
            if name == "u3":
                inv.append(("u3", t, (-prm[0], -prm[2], -prm[1]), c, ng))
            else:
                inv.append((name, t, prm, c, ng))
        sim.run(inv)
        p0 = sim.probability_total()
        assert abs(p0 - 1) < 1e-12
        out = sim.sample(1024, seed=2)
        assert set(np.unique(out).tolist()) <= {0, (1 << n) - 1}
