"""Seeded circuit generators for the benchmark workloads.

A circuit is a list of ``(name, targets, params, ctrls, negctrls)`` tuples in the argument
order of the C API.  All randomness comes from SplitMix64 (Steele, Lea & Flood, OOPSLA
2014) so that the same seed gives the same circuit in every process and language:
``u = (splitmix64(state += 0x9E3779B97F4A7C15) >> 11) * 2**-53``.

Workloads:

* ``ghz_u3_cz`` -- H(q0), CNOT q0->qi for i = 1..n-1 (the GHZ construction of the
  reference example examples/standalone/ghz_circuit.c:36-41), then ``layers`` layers of
  U3(theta, phi, lambda ~ U[0, 2pi)) on every qubit followed by CZ on a random perfect
  matching (Fisher-Yates shuffle, consecutive pairs).
* ``brick`` -- ``depth`` layers of U3 on every qubit, then CZ on the even pairs
  (0,1),(2,3),... in even layers and the odd pairs (1,2),(3,4),... in odd layers; plus
  ``mcx`` multi-controlled X gates with 3 controls inserted after random layers, each with
  at least one operand among the top 6 qubits (high-stride qubits).
* ``variational`` -- ``depth`` layers of RY and RZ (U[-pi, pi]) on every qubit plus a
  CNOT ladder q_i -> q_{i+1}.
* ``qft`` -- H plus controlled-phase ladder plus the final swaps.
* ``qaoa_ring`` -- QAOA rounds of RZZ on a ring and RX mixers (planner workload).
* ``narrow_random`` -- a deep random circuit on the top qubits only (planner workload).
"""
from __future__ import annotations

import math

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(z: int) -> int:
    z &= MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK

    def next_u64(self) -> int:
        self.state = (self.state + GOLDEN) & MASK
        return splitmix64(self.state)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (2.0 ** -53)

    def randbelow(self, k: int) -> int:
        return min(int(self.uniform() * k), k - 1)


def uniform_stream(seed: int, count: int):
    """The sampling uniforms of qcs_ext_sample: u_j = splitmix64(seed + (j+1)*gamma)."""
    return [(splitmix64(seed + (j + 1) * GOLDEN) >> 11) * (2.0 ** -53) for j in range(count)]


def _u3_layer(rng, n):
    out = []
    for q in range(n):
        th, ph, la = (2 * math.pi * rng.uniform() for _ in range(3))
        out.append(("u3", (q,), (th, ph, la), (), ()))
    return out


def ghz_u3_cz(n: int, layers: int, seed: int):
    rng = SplitMix64(seed)
    circ = [("h", (0,), (), (), ())]
    for q in range(1, n):
        circ.append(("x", (q,), (), (0,), ()))
    for _ in range(layers):
        circ += _u3_layer(rng, n)
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.randbelow(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for k in range(0, n - 1, 2):
            a, b = perm[k], perm[k + 1]
            circ.append(("z", (b,), (), (a,), ()))
    return circ


def brick(n: int, depth: int, seed: int, mcx: int = 0):
    rng = SplitMix64(seed)
    after = sorted(rng.randbelow(depth) for _ in range(mcx))
    circ = []
    for layer in range(depth):
        circ += _u3_layer(rng, n)
        start = layer % 2
        for a in range(start, n - 1, 2):
            circ.append(("z", (a + 1,), (), (a,), ()))
        while after and after[0] == layer:
            after.pop(0)
            high = n - 1 - rng.randbelow(min(6, n))
            ops = [high]
            while len(ops) < 4:
                q = rng.randbelow(n)
                if q not in ops:
                    ops.append(q)
            # shuffle which operand is the target
            t = ops.pop(rng.randbelow(4))
            circ.append(("x", (t,), (), tuple(ops), ()))
    return circ


def variational(n: int, depth: int, seed: int):
    rng = SplitMix64(seed)
    circ = []
    for _ in range(depth):
        for q in range(n):
            circ.append(("ry", (q,), ((2 * rng.uniform() - 1) * math.pi,), (), ()))
            circ.append(("rz", (q,), ((2 * rng.uniform() - 1) * math.pi,), (), ()))
        for q in range(n - 1):
            circ.append(("x", (q + 1,), (), (q,), ()))
    return circ


def variational_batch(count: int, seed: int, nmin: int = 12, nmax: int = 22, dmin: int = 4, dmax: int = 16):
    """Batch workload: `count` variational circuits, qubits ~ U{nmin..nmax}, depth ~
    U{dmin..dmax}; each gate as (name, target, params, control) for pyqcs.Batch."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = nmin + rng.randbelow(nmax - nmin + 1)
        depth = dmin + rng.randbelow(dmax - dmin + 1)
        gates = []
        for name, t, prm, c, _ng in variational(n, depth, rng.next_u64()):
            gates.append((name, t[0], prm, c[0] if c else -1))
        out.append((n, gates))
    return out


def batch_to_circuit(gates):
    """(name, target, params, control) gates -> qcs circuit tuples."""
    return [(g[0], (g[1],), tuple(g[2]), (g[3],) if g[3] >= 0 else (), ()) for g in gates]


def qft(n: int):
    circ = []
    for j in range(n - 1, -1, -1):
        circ.append(("h", (j,), (), (), ()))
        for k in range(j - 1, -1, -1):
            circ.append(("p", (j,), (math.pi / (1 << (j - k)),), (k,), ()))
    for i in range(n // 2):
        circ.append(("swap", (i, n - 1 - i), (), (), ()))
    return circ


def qaoa_ring(n: int, p: int, seed: int):
    """QAOA on a ring: H on every qubit, then p rounds of RZZ(gamma) on the pairs (q, q+1
    mod n) and RX(beta) on every qubit, gamma and beta ~ U[0, pi) per round."""
    rng = SplitMix64(seed)
    circ = [("h", (q,), (), (), ()) for q in range(n)]
    for _ in range(p):
        gamma, beta = math.pi * rng.uniform(), math.pi * rng.uniform()
        circ += [("rzz", (q, (q + 1) % n), (gamma,), (), ()) for q in range(n)]
        circ += [("rx", (q,), (beta,), (), ()) for q in range(n)]
    return circ


def narrow_random(n: int, width: int, count: int, seed: int):
    """`count` random gates on the top `width` qubits only: U3 (angles ~ U[0, 2pi)) or CX
    on a random ordered pair, each with probability 1/2 (a deep circuit on a register at the
    high end of the index, the case where a wide tile of high qubits serves a whole flush)."""
    rng = SplitMix64(seed)
    qs = list(range(n - width, n))
    circ = []
    for _ in range(count):
        if rng.uniform() < 0.5:
            q = qs[rng.randbelow(width)]
            circ.append(("u3", (q,), tuple(2 * math.pi * rng.uniform() for _ in range(3)), (), ()))
        else:
            a = qs[rng.randbelow(width)]
            b = a
            while b == a:
                b = qs[rng.randbelow(width)]
            circ.append(("x", (b,), (), (a,), ()))
    return circ


def random_two_qubit(n: int, count: int, seed: int):
    """U3 (40 %), CX (30 %) and controlled phase (30 %) on random qubits (pairs anywhere)."""
    rng = SplitMix64(seed)
    circ = []
    for _ in range(count):
        r = rng.uniform()
        a = rng.randbelow(n)
        b = a
        while b == a:
            b = rng.randbelow(n)
        if r < 0.4:
            circ.append(("u3", (a,), tuple(2 * math.pi * rng.uniform() for _ in range(3)), (), ()))
        elif r < 0.7:
            circ.append(("x", (b,), (), (a,), ()))
        else:
            circ.append(("p", (b,), (2 * math.pi * rng.uniform(),), (a,), ()))
    return circ


def diagonal_heavy(n: int, count: int, seed: int):
    """H on every qubit, then controlled phases (45 %), RZ (45 %) and H (10 %) on random
    qubits: stages that are mostly phase tables."""
    rng = SplitMix64(seed)
    circ = [("h", (q,), (), (), ()) for q in range(n)]
    for _ in range(count):
        r = rng.uniform()
        a = rng.randbelow(n)
        b = a
        while b == a:
            b = rng.randbelow(n)
        if r < 0.45:
            circ.append(("p", (b,), (2 * math.pi * rng.uniform(),), (a,), ()))
        elif r < 0.9:
            circ.append(("rz", (a,), (2 * math.pi * rng.uniform(),), (), ()))
        else:
            circ.append(("h", (a,), (), (), ()))
    return circ


def planner_training_set(n: int, seed: int):
    """Circuits whose queued gates the planner's plan recorder times (one of each family,
    sizes and depths varied by the seed): QFT on a brick layer, brick, QAOA ring, a narrow
    register, random two-qubit gates, diagonal-heavy."""
    return {
        "qft": brick(n, 1, seed=seed) + qft(n),
        "brick": brick(n, 8 + seed % 16, seed=seed),
        "qaoa": qaoa_ring(n, 1 + seed % 4, seed),
        "narrow": narrow_random(n, 6 + seed % 8, 100 + 40 * (seed % 6), seed),
        "random2q": random_two_qubit(n, 200 + 50 * (seed % 5), seed),
        "diagonal": diagonal_heavy(n, 200 + 50 * (seed % 5), seed),
    }


def random_all_gates(n: int, count: int, seed: int, max_ctrls: int = 2):
    """Random circuit over the whole gate set, random positive/negative controls."""
    from . import GATE_PARAMS, GATE_TARGETS

    names = sorted(GATE_TARGETS)
    rng = SplitMix64(seed)
    circ = []
    while len(circ) < count:
        name = names[rng.randbelow(len(names))]
        t = GATE_TARGETS[name]
        nc = rng.randbelow(max_ctrls + 1)
        if t + nc > n:
            continue
        qs = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.randbelow(i + 1)
            qs[i], qs[j] = qs[j], qs[i]
        targets = tuple(qs[:t])
        ctl = qs[t:t + nc]
        pos = tuple(c for c in ctl if rng.uniform() < 0.6)
        neg = tuple(c for c in ctl if c not in pos)
        npar = 1 if name == "global_phase" else GATE_PARAMS.get(name, 0)
        params = tuple(2 * math.pi * rng.uniform() - math.pi for _ in range(npar))
        circ.append((name, targets, params, pos, neg))
    return circ


def algorithmic_bytes(n: int, op, amp_bytes: int = 16) -> float:
    """Bytes a gate must move: 2 * S * A, A = amplitudes its matrix can change."""
    name, targets, params, ctrls, negctrls = op
    c = len(ctrls) + len(negctrls)
    if name in ("z", "s", "sdg", "t", "tdg", "p", "u1", "swap", "iswap", "xx_plus_yy", "xx_minus_yy"):
        a = 2 ** (n - c - 1)
    elif name == "dcx":
        a = 3 * 2 ** (n - c - 2)
    elif name == "rccx":
        a = 3 * 2 ** (n - c - 3)
    elif name == "rcccx":
        a = 2 ** (n - c - 2)
    elif name == "id":
        a = 0
    else:
        a = 2 ** (n - c)
    return 2.0 * amp_bytes * a
