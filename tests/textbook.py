"""Textbook dense model of the gate set, used to pin the oracle.

Matrices are the published Qiskit conventions (little-endian: target j of a gate is bit j
of the gate's local basis index), written from the textbook definitions, not from the
reference code.  A gate is applied by tensor contraction over the target axes of the
state tensor, and controls are applied as a projector: the output keeps the old amplitude
wherever a control is not satisfied.  Nothing here shares code with the oracle's
index-insertion kernel, so an error in either shows up as a mismatch.
"""
from __future__ import annotations

import numpy as np

SQ = 1 / np.sqrt(2)


def _c_s(theta):
    return np.cos(theta / 2), np.sin(theta / 2)


def matrix(name, params=()):
    p = list(params)
    e = lambda a: np.exp(1j * a)  # noqa: E731
    if name == "global_phase":
        return np.array([[e(p[0])]])
    if name == "h":
        return SQ * np.array([[1, 1], [1, -1]], dtype=complex)
    if name == "x":
        return np.array([[0, 1], [1, 0]], dtype=complex)
    if name == "y":
        return np.array([[0, -1j], [1j, 0]])
    if name == "z":
        return np.diag([1, -1]).astype(complex)
    if name == "s":
        return np.diag([1, 1j])
    if name == "sdg":
        return np.diag([1, -1j])
    if name == "t":
        return np.diag([1, e(np.pi / 4)])
    if name == "tdg":
        return np.diag([1, e(-np.pi / 4)])
    if name == "sx":
        return 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]])
    if name == "sxdg":
        return 0.5 * np.array([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]])
    if name == "id":
        return np.eye(2, dtype=complex)
    if name == "rx":
        c, s = _c_s(p[0])
        return np.array([[c, -1j * s], [-1j * s, c]])
    if name == "ry":
        c, s = _c_s(p[0])
        return np.array([[c, -s], [s, c]], dtype=complex)
    if name == "rz":
        return np.diag([e(-p[0] / 2), e(p[0] / 2)])
    if name in ("u3", "u"):
        c, s = _c_s(p[0])
        return np.array([[c, -e(p[2]) * s], [e(p[1]) * s, e(p[1] + p[2]) * c]])
    if name == "u4":
        return e(p[3]) * matrix("u3", p[:3])
    if name == "u2":
        return matrix("u3", [np.pi / 2, p[0], p[1]])
    if name in ("u1", "p"):
        return np.diag([1, e(p[0])])
    if name == "r":
        c, s = _c_s(p[0])
        return np.array([[c, -1j * e(-p[1]) * s], [-1j * e(p[1]) * s, c]])
    # two-qubit gates: local index = b0 + 2*b1 (b0 = first target)
    X = matrix("x")
    Y = matrix("y")
    Z = matrix("z")
    I2 = np.eye(2)
    if name == "swap":
        return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)
    if name == "iswap":
        return np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])
    if name == "rzz":
        return np.diag([e(-p[0] / 2), e(p[0] / 2), e(p[0] / 2), e(-p[0] / 2)])
    if name == "rxx":
        c, s = _c_s(p[0])
        return c * np.eye(4) - 1j * s * np.kron(X, X)
    if name == "ryy":
        c, s = _c_s(p[0])
        return c * np.eye(4) - 1j * s * np.kron(Y, Y)
    if name == "rzx":
        # exp(-i theta/2 Z_{t0} X_{t1}); kron puts the higher qubit (t1) first
        c, s = _c_s(p[0])
        return c * np.eye(4) - 1j * s * np.kron(X, Z)
    if name == "xx_plus_yy":
        c, s = _c_s(p[0])
        m = np.eye(4, dtype=complex)
        m[1, 1] = c
        m[1, 2] = -1j * s * e(-p[1])
        m[2, 1] = -1j * s * e(p[1])
        m[2, 2] = c
        return m
    if name == "xx_minus_yy":
        c, s = _c_s(p[0])
        m = np.eye(4, dtype=complex)
        m[0, 0] = c
        m[0, 3] = -1j * s * e(-p[1])
        m[3, 0] = -1j * s * e(p[1])
        m[3, 3] = c
        return m
    if name == "dcx":
        # CX(control t0 -> target t1) followed by CX(control t1 -> target t0)
        cx01 = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]], dtype=complex)
        cx10 = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
        return cx10 @ cx01
    if name == "ecr":
        return SQ * np.array([[0, 1, 0, 1j], [1, 0, -1j, 0], [0, 1j, 0, 1], [-1j, 0, 1, 0]])
    if name == "rccx":
        m = np.eye(8, dtype=complex)
        m[3, 3] = 0
        m[7, 7] = 0
        m[3, 7] = -1j
        m[7, 3] = 1j
        m[5, 5] = -1
        return m
    if name == "rcccx":
        m = np.eye(16, dtype=complex)
        m[3, 3] = 1j
        m[11, 11] = -1j
        m[7, 7] = 0
        m[15, 15] = 0
        m[7, 15] = 1
        m[15, 7] = -1
        return m
    del I2
    raise KeyError(name)


def apply(psi, n, name, targets, params=(), ctrls=(), negctrls=()):
    """Return the state after the gate; psi is a length-2^n complex vector."""
    G = matrix(name, params)
    t = len(targets)
    psi = np.asarray(psi, dtype=complex)
    if t == 0:
        out = G[0, 0] * psi
    else:
        # tensor axis a <-> qubit n-1-a (C order: first axis is the most significant bit)
        tens = psi.reshape((2,) * n)
        Gt = G.reshape((2,) * (2 * t))  # out axes (b_{t-1},...,b_0), in axes likewise
        in_axes = list(range(t, 2 * t))
        # Gt in-axis t+k corresponds to bit t-1-k, i.e. target[t-1-k]
        state_axes = [n - 1 - targets[t - 1 - k] for k in range(t)]
        res = np.tensordot(Gt, tens, axes=(in_axes, state_axes))
        # res axes: G out axes (for targets[t-1],...,targets[0]) then the remaining state axes
        remaining = [a for a in range(n) if a not in state_axes]
        order = state_axes + remaining
        inv = np.argsort(order)
        out = np.transpose(res, inv).reshape(-1)
    idx = np.arange(1 << n)
    ok = np.ones(1 << n, dtype=bool)
    for c in ctrls:
        ok &= ((idx >> c) & 1) == 1
    for c in negctrls:
        ok &= ((idx >> c) & 1) == 0
    return np.where(ok, out, psi)


def run(n, circuit, psi=None):
    if psi is None:
        psi = np.zeros(1 << n, dtype=complex)
        psi[0] = 1
    for op in circuit:
        name, targets, params, ctrls, negctrls = (list(op) + [(), (), ()])[:5]
        psi = apply(psi, n, name, list(targets), params, list(ctrls), list(negctrls))
    return psi
