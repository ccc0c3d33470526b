"""Learned fusion planner: the cost-model MLP and its use by the fused engine.

The MLP's forward pass and gradients are checked three ways: an independent numpy
implementation (pinned by central finite differences in float64), the oracle library's CPU
loops, and the GPU kernels of libqcs.so.
"""
import numpy as np
import pytest

import pyqcs
from pyqcs import circuits
from conftest import fidelity

H = 64


def unpack(p, F):
    o = 0
    W1 = p[o:o + H * F].reshape(H, F); o += H * F
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def np_forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1])
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ w3 + b3, h1, h2


def np_loss(p, X, T):
    y, _, _ = np_forward(p, X)
    return np.mean((y - T) ** 2)


def np_grad(p, X, T):
    F = X.shape[1]
    W1, b1, W2, b2, w3, b3 = unpack(p, F)
    y, h1, h2 = np_forward(p, X)
    B = X.shape[0]
    dy = 2 * (y - T) / B
    d2 = np.outer(dy, w3) * (h2 > 0)
    d1 = (d2 @ W2) * (h1 > 0)
    return np.concatenate([(d1.T @ X).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), h2.T @ dy, [dy.sum()]])


def data(F=16, B=48, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, size=(B, F)).astype(np.float32)
    T = (np.sin(3 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3]).astype(np.float32)
    return X, T


def test_numpy_gradient_pinned_by_finite_differences():
    F = 5
    X, T = data(F, 20, 1)
    rng = np.random.default_rng(2)
    p = rng.normal(0, 0.4, size=H * F + 3 * H + H * H + 1)
    X, T = X.astype(np.float64), T.astype(np.float64)
    g = np_grad(p, X, T)
    for i in rng.choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[i] = 1e-6
        fd = (np_loss(p + e, X, T) - np_loss(p - e, X, T)) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i])


def test_oracle_mlp_matches_numpy(oracle):
    X, T = data()
    m = pyqcs.Mlp(oracle, 16, seed=7)
    p = m.params()
    assert p.size == H * 16 + 3 * H + H * H + 1
    y = m.predict(X)
    want, _, _ = np_forward(p.astype(np.float64), X.astype(np.float64))
    assert np.max(np.abs(y - want)) < 1e-5
    g, loss = m.gradients(X, T)
    gw = np_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert np.max(np.abs(g - gw)) <= 1e-5 + 1e-4 * np.max(np.abs(gw))
    assert abs(loss - np_loss(p.astype(np.float64), X, T)) < 1e-5
    l0 = loss
    l1 = m.train(X, T, epochs=200, batch=16, lr=3e-3, seed=3)
    assert l1 < 0.3 * l0


@pytest.mark.gpu
def test_gpu_mlp_matches_oracle(gpu, oracle):
    X, T = data(B=300, seed=4)
    mg = pyqcs.Mlp(gpu, 16, seed=11)
    mo = pyqcs.Mlp(oracle, 16, seed=11)
    p = mg.params()
    assert np.array_equal(p, mo.params())  # same SplitMix64 initialisation
    assert np.max(np.abs(mg.predict(X) - mo.predict(X))) < 1e-5
    gg, lg = mg.gradients(X, T)
    go, lo = mo.gradients(X, T)
    assert np.max(np.abs(gg - go)) <= 1e-6 + 1e-4 * np.max(np.abs(go))
    assert abs(lg - lo) < 1e-5
    l1 = mg.train(X, T, epochs=150, batch=32, lr=3e-3, seed=5)
    assert l1 < 0.3 * lg


def test_feature_count(gpu, oracle):
    assert pyqcs.planner_num_features(oracle) == pyqcs.planner_num_features(gpu) == 20


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [(1, 1e-6), (2, 1e-12)])
def test_planner_record_train_and_plan(gpu, oracle, precision, tol):
    n = 20
    F = pyqcs.planner_num_features(gpu)
    with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=5) as sim:
        X, y = pyqcs.planner_record(sim, 400, seed=9)
    assert X.shape == (400, F) and np.all(np.isfinite(y))
    assert np.all(X[:, F - 1] == 1) and np.all(X[:, 0] == np.float32(n / 32))
    stages = X[:, 4] == 1
    assert 0 < stages.sum() < len(stages)  # both launch kinds recorded
    model = pyqcs.Mlp(gpu, F, seed=3)
    loss = model.train(X, y, epochs=100, batch=64, lr=3e-3, seed=1)
    assert loss < np.var(y)  # better than predicting the mean
    circ = circuits.qft(n) + circuits.brick(n, 4, seed=3)
    states = []
    for use in (False, True):
        with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=5) as sim:
            if use:
                pyqcs.set_planner(sim, model)
            sim.run(circ)
            states.append(sim.statevector())
            st = pyqcs.planner_stats(sim)
            if use:
                assert st["flushes"] >= 1 and st["pred_chosen_us"] <= st["pred_greedy_us"] * 1.0001
    with pyqcs.Simulator(oracle, n) as o:
        o.run(circ)
        want = o.statevector()
    for s in states:
        assert fidelity(s, want) >= 1 - tol


def _median_run(sim, circ, reps=5):
    import time

    ts = []
    for _ in range(reps):
        sim.call("set_zero_state")
        sim.flush()
        sim.norm()
        t0 = time.perf_counter()
        sim.run(circ)
        sim.flush()
        sim.norm()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


@pytest.mark.gpu
def test_learned_planner_measured_gain(gpu):
    # A cost model trained on the GPU (random launches at 26-30 qubits, and the launches of
    # every candidate stage plan on training circuits at 26 and 28 qubits, other circuits
    # than these) chooses per flush among the stage plans: tile bits 11-14, 4- or 5-qubit
    # blocks, grown or window tiles.  On a QAOA ring and a deep circuit on the top 10
    # qubits the default plan (2^13 tiles, 5-qubit blocks) is not the fastest; the learned
    # choice must be measurably faster on at least one of them (median of 5 runs) and
    # within noise of the default on both, with the same output state.
    n = 30
    model, info = pyqcs.train_cost_model(gpu, 3, 5, (26, 28, 30), 2000, (26, 28))
    print(f"cost model: {info}", flush=True)
    work = {"qaoa": circuits.qaoa_ring(n, 3, 7), "narrow": circuits.narrow_random(n, 10, 300, 8)}
    ratios = {}
    try:
        with pyqcs.Simulator(gpu, n, seed=1, precision=3, block_qubits=5) as d, \
                pyqcs.Simulator(gpu, n, seed=1, precision=3, block_qubits=5) as learned:
            pyqcs.set_planner(learned, model)
            for name, circ in work.items():
                t_default = _median_run(d, circ)
                t_learned = _median_run(learned, circ)
                choice = pyqcs.planner_last_choice(learned)
                ov, md = learned.overlap(d)
                fid = abs(ov) ** 2 / (learned.norm() * d.norm())
                ratios[name] = t_learned / t_default
                print(f"{name}: default {t_default * 1e3:.1f} ms, learned {t_learned * 1e3:.1f} ms "
                      f"(plan {choice}), fidelity between them {fid:.12f}", flush=True)
                assert fid >= 1 - 1e-6 and md <= 1e-5
            pyqcs.set_planner(learned, None)
    finally:
        model.close()
    assert max(ratios.values()) <= 1.05
    assert min(ratios.values()) <= 0.95


@pytest.mark.gpu
def test_planner_check_keeps_the_faster_plan(gpu, oracle):
    # A hand-set model whose prediction falls with the tile bits (feature 17 = m / 16, through
    # one hidden unit of each layer: log2 us = b3 - 32 * m / 16; block launches, with feature
    # 17 = 0, are priced at b3) rates a stage launch on 2^14 tiles 4x cheaper than one on
    # 2^13, so it overrides the default plan (2^13 tiles, 5-qubit blocks, fewer-stage tile
    # choice) with a 2^14-tile plan.  With the measured
    # check on, the first run of the circuit takes the model's plan, the second the default
    # plan, and the later runs the faster of the two; with it off, every run takes the
    # model's plan.  Every run's state matches the oracle.
    n = 20
    F = pyqcs.planner_num_features(gpu)
    model = pyqcs.Mlp(gpu, F, seed=1)
    p = np.zeros(model.P, dtype=np.float32)
    p[17] = 1.0                      # W1[0][17]
    p[H * F + H] = 1.0               # W2[0][0]
    p[H * F + H + H * H + H] = -32.0  # w3[0]
    p[-1] = np.log2(100.0) + 28.0    # b3
    model.set_params(p)
    X = np.zeros((3, F), dtype=np.float32)
    X[:, 17] = [0.0, 13 / 16, 14 / 16]
    assert np.allclose(model.predict(X), [np.log2(100.0) + 28, np.log2(100.0) + 2, np.log2(100.0)], atol=1e-4)
    circ = circuits.brick(n, 6, seed=11)
    with pyqcs.Simulator(oracle, n) as o:
        o.run(circ)
        want = o.statevector()
    try:
        for check in (True, False):
            with pyqcs.Simulator(gpu, n, seed=1, precision=3, block_qubits=5) as sim:
                pyqcs.set_planner(sim, model)
                pyqcs.planner_set_check(sim, check)
                choices = []
                for r in range(4):
                    sim.call("set_zero_state")
                    sim.run(circ)
                    assert fidelity(sim.statevector(), want) >= 1 - 1e-6
                    choices.append(tuple(pyqcs.planner_last_choice(sim)))
                st = pyqcs.planner_stats(sim)
                pyqcs.set_planner(sim, None)
            print(f"check {check}: choices {choices}, stats {st}", flush=True)
            assert st["flushes"] == 4
            if check:
                learned, default = choices[0], choices[1]
                assert learned[0] == 14 and default[:3] == (13, 5, 2)
                assert choices[2] == choices[3] and choices[2] in (learned, default)
                assert st["learned_choices"] == (3 if choices[2] == learned else 1)
            else:
                assert len(set(choices)) == 1 and choices[0][0] == 14
                assert st["learned_choices"] == 4
    finally:
        model.close()


@pytest.mark.gpu
def test_planner_check_per_structure_and_reset(gpu):
    # The measured check keeps one record per flush structure (gate names, kinds and
    # qubits; parameters ignored): two circuits of different structure run in alternation
    # each go through model plan -> default plan -> decision on their own; a circuit that
    # differs from one of them only in its angles takes that structure's decision at once;
    # and a change of the stage tile bits forgets the records (the next run is the model's).
    n = 20
    F = pyqcs.planner_num_features(gpu)
    model = pyqcs.Mlp(gpu, F, seed=1)
    p = np.zeros(model.P, dtype=np.float32)
    p[17] = 1.0
    p[H * F + H] = 1.0
    p[H * F + H + H * H + H] = -32.0
    p[-1] = np.log2(100.0) + 28.0
    model.set_params(p)
    circs = [circuits.brick(n, 6, seed=11), circuits.brick(n, 5, seed=12)]
    same_structure = circuits.brick(n, 6, seed=13)  # circs[0] with other angles

    def run(sim, circ):
        sim.call("set_zero_state")
        sim.run(circ)
        sim.flush()
        return tuple(pyqcs.planner_last_choice(sim))

    try:
        with pyqcs.Simulator(gpu, n, seed=1, precision=3, block_qubits=5) as sim:
            pyqcs.set_planner(sim, model)
            seq = {0: [], 1: []}
            for r in range(3):
                for i, circ in enumerate(circs):
                    seq[i].append(run(sim, circ))
            print(f"alternating: {seq}", flush=True)
            for i in (0, 1):
                learned, default, decided = seq[i]
                assert learned != default and default[:3] == (13, 5, 2) and decided in (learned, default)
            assert run(sim, same_structure) == seq[0][2]
            sim.set_stage_bits(12)
            sim.set_stage_bits(13)
            assert run(sim, circs[0]) == seq[0][0]  # checked afresh
            pyqcs.set_planner(sim, None)
    finally:
        model.close()
