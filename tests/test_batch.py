"""Batching form: many circuits, <Z_q> outputs; oracle = plain method per circuit."""
import math

import numpy as np
import pytest

import pyqcs
from pyqcs import circuits
import textbook


def expz_textbook(n, gates):
    psi = textbook.run(n, circuits.batch_to_circuit(gates))
    p = np.abs(psi) ** 2
    idx = np.arange(1 << n)
    return np.array([np.sum(np.where((idx >> q) & 1, -p, p)) for q in range(n)])


def test_oracle_batch_matches_textbook(oracle):
    work = circuits.variational_batch(6, seed=3, nmin=3, nmax=7, dmin=1, dmax=3)
    work.append((4, [("h", 0, (), -1), ("x", 1, (), 0), ("u3", 3, (0.3, 0.2, 0.1), 2), ("rx", 2, (1.1,), -1)]))
    b = pyqcs.Batch(oracle)
    for n, g in work:
        b.add(n, g)
    res, _ = b.run()
    for (n, g), z in zip(work, res):
        assert np.max(np.abs(z - expz_textbook(n, g))) < 1e-13
    assert b.info()["circuits"] == len(work)
    again, _ = b.run_plain()  # the oracle's batch run is the plain method already
    assert all(np.array_equal(x, y) for x, y in zip(res, again))


def test_batch_rejects_bad_input(oracle):
    b = pyqcs.Batch(oracle)
    with pytest.raises(pyqcs.QcsError):
        b.add(3, [("swap", 0, (), -1)])  # two-target gate
    with pytest.raises(pyqcs.QcsError):
        b.add(3, [("h", 3, (), -1)])  # target out of range
    with pytest.raises(pyqcs.QcsError):
        b.add(3, [("x", 1, (), 1)])  # control == target
    b.set_precision(3)
    b.set_precision(1)
    for mode in (0, 2, 4):
        with pytest.raises(pyqcs.QcsError):
            b.set_precision(mode)


@pytest.mark.gpu
def test_gpu_batch_vs_oracle(gpu, oracle):
    work = circuits.variational_batch(40, seed=5, nmin=4, nmax=17, dmin=2, dmax=6)
    work.append((9, [("h", q, (), -1) for q in range(9)] + [("z", 4, (), 3), ("s", 2, (), -1), ("t", 7, (), 6)]))
    ref = pyqcs.Batch(oracle)
    for n, g in work:
        ref.add(n, g)
    want, _ = ref.run()
    for tuning in [(14, 512, 5, 4, 12), (12, 256, 4, 1, 0), (13, 1024, 5, 2, 12), (0, 256, 5, 4, 12), (0, 512, 4, 3, 0)]:
        b = pyqcs.Batch(gpu)
        for n, g in work:
            b.add(n, g)
        b.set_tuning(*tuning)
        got, secs = b.run()
        err = max(np.max(np.abs(x - y)) for x, y in zip(got, want))
        print(f"tuning {tuning}: max |d<Z>| {err:.2e} device {secs * 1e3:.2f} ms nodes {b.info()['graph_nodes']}", flush=True)
        assert err < 2e-5, tuning
        # graph replays give identical bits
        for _ in range(2):
            again, _ = b.run()
            assert all(np.array_equal(x, y) for x, y in zip(got, again))
        b.close()


@pytest.mark.gpu
def test_gpu_batch_stage_rounds(gpu, oracle):
    # circuits of 15-18 qubits with different depths: with 2^11-amplitude tiles all of them
    # run in stage rounds (one launch per round across circuits, later rounds with fewer
    # circuits); a gate-free 16-qubit circuit and small LDS circuits ride along
    work = circuits.variational_batch(10, seed=21, nmin=15, nmax=18, dmin=1, dmax=7)
    work += [(16, []), (6, [("h", 0, (), -1), ("x", 5, (), 0)])]
    ref = pyqcs.Batch(oracle)
    for n, g in work:
        ref.add(n, g)
    want, _ = ref.run()
    results = {}
    for tuning in [(14, 512, 5, 2, 11), (14, 512, 4, 4, 11), (14, 512, 5, 1, 0)]:
        b = pyqcs.Batch(gpu)
        for n, g in work:
            b.add(n, g)
        b.set_tuning(*tuning)
        got, secs = b.run()
        direct, _ = b.run_direct()
        err = max(np.max(np.abs(x - y)) for x, y in zip(got, want))
        print(f"tuning {tuning}: max |d<Z>| {err:.2e} device {secs * 1e3:.2f} ms nodes {b.info()['graph_nodes']}", flush=True)
        assert err < 2e-5, tuning
        assert all(np.array_equal(x, y) for x, y in zip(got, direct))
        assert np.array_equal(got[-2], np.ones(16))  # |0...0>: <Z_q> = 1 exactly
        results[tuning] = got
        b.close()
    # rounds and per-block launches agree to complex64 rounding
    a, c = results[(14, 512, 5, 2, 11)], results[(14, 512, 5, 1, 0)]
    assert max(np.max(np.abs(x - y)) for x, y in zip(a, c)) < 2e-5


@pytest.mark.gpu
def test_gpu_batch_bf16x3(gpu, oracle):
    # BF16x3 split-precision MFMA for the circuits in HBM: stage rounds (2^11 and 2^12 tiles)
    # and per-block launches (stage_bits 0), against the oracle's FP64 plain method; the
    # LDS-resident circuits (<= 12 or 14 qubits) take the same path in both modes
    work = circuits.variational_batch(12, seed=31, nmin=10, nmax=18, dmin=1, dmax=6)
    ref = pyqcs.Batch(oracle)
    for n, g in work:
        ref.add(n, g)
    want, _ = ref.run()
    for tuning in [(14, 512, 4, 2, 11), (14, 512, 5, 1, 12), (12, 256, 4, 2, 0), (12, 256, 5, 1, 0)]:
        got = {}
        for mode in (1, 3):
            b = pyqcs.Batch(gpu)
            for n, g in work:
                b.add(n, g)
            b.set_tuning(*tuning)
            b.set_precision(mode)
            got[mode], _ = b.run()
            again, _ = b.run()
            direct, _ = b.run_direct()
            assert all(np.array_equal(x, y) for x, y in zip(got[mode], again))
            assert all(np.array_equal(x, y) for x, y in zip(got[mode], direct))
            b.close()
        err = {m: max(np.max(np.abs(x - y)) for x, y in zip(got[m], want)) for m in (1, 3)}
        print(f"tuning {tuning}: max |d<Z>| FP32 {err[1]:.2e} BF16x3 {err[3]:.2e}", flush=True)
        assert err[1] < 2e-5 and err[3] < 5e-5, tuning
        # the modes run different products on the circuits in HBM
        assert any(not np.array_equal(x, y) for x, y in zip(got[1], got[3])), tuning


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [1, 3])
def test_gpu_batch_large_circuits_vs_oracle(gpu, oracle, mode):
    # the upper end of the benchmark batch, 19-22 qubits (complex64 states in HBM, stage
    # rounds), beside small LDS-resident circuits, against the oracle's FP64 plain method;
    # also the GPU plain loop (complex128, one kernel per gate) at those sizes
    work = circuits.variational_batch(5, seed=41, nmin=19, nmax=22, dmin=2, dmax=5)
    work += circuits.variational_batch(4, seed=42, nmin=6, nmax=12, dmin=2, dmax=5)
    assert max(n for n, _ in work) >= 21
    ref = pyqcs.Batch(oracle)
    b = pyqcs.Batch(gpu)
    for n, g in work:
        ref.add(n, g)
        b.add(n, g)
    want, _ = ref.run()
    b.set_precision(mode)
    got, secs = b.run()
    err = max(np.max(np.abs(x - y)) for x, y in zip(got, want))
    print(f"19-22 qubit batch mode {mode}: max |d<Z>| {err:.2e} device {secs * 1e3:.2f} ms", flush=True)
    assert err < (2e-5 if mode == 1 else 5e-5)
    if mode == 1:
        plain, _ = b.run_plain()
        assert max(np.max(np.abs(x - y)) for x, y in zip(plain, want)) < 1e-12
    b.close()


@pytest.mark.gpu
def test_gpu_batch_traffic(gpu, oracle):
    # a gate-free 16-qubit circuit in HBM: one fill (8 B per amplitude) and one <Z>
    # reduction (8 B), 3 launches; an LDS-resident circuit: no HBM bytes, 2 launches
    b = pyqcs.Batch(gpu)
    b.add(16, [])
    b.add(5, [("h", 0, (), -1)])
    assert b.traffic() == {"hbm_bytes": 16.0 * 2 ** 16, "kernel_launches": 5}
    b.close()
    # staged circuits: one sweep (16 B per amplitude) per stage round, less the first round's
    # reads (it starts from |0...0> without loading), plus the reduction's read (8 B)
    b = pyqcs.Batch(gpu)
    b.add(16, [("ry", q, (0.1 * q,), -1) for q in range(16)])
    b.set_tuning(14, 512, 5, 1, 11)
    t = b.traffic()
    sweeps = t["hbm_bytes"] / (16.0 * 2 ** 16)
    assert sweeps == int(sweeps) and sweeps >= 1
    assert t["kernel_launches"] == 2 + int(sweeps)  # rounds, reduction pair
    b.close()
    o = pyqcs.Batch(oracle)
    o.add(4, [])
    assert o.traffic() == {"hbm_bytes": 0.0, "kernel_launches": 0}


@pytest.mark.gpu
def test_gpu_batch_autotune(gpu):
    work = circuits.variational_batch(24, seed=8, nmin=10, nmax=18, dmin=2, dmax=5)
    b = pyqcs.Batch(gpu)
    for n, g in work:
        b.add(n, g)
    before, _ = b.run()
    t = b.autotune()
    assert t["lds_max_qubits"] in (12, 13, 14) and t["block_size"] in (256, 512, 1024)
    assert t["block_qubits"] in (4, 5) and t["streams"] in (1, 2, 4)
    assert t["stage_bits"] in (0, 11, 12, 13)
    after, _ = b.run()
    assert max(np.max(np.abs(x - y)) for x, y in zip(before, after)) < 2e-5
    # the search within BF16x3 products
    b.set_precision(3)
    t3 = b.autotune()
    assert t3["block_qubits"] in (4, 5) and t3["stage_bits"] in (0, 11, 12, 13)
    fast, _ = b.run()
    assert max(np.max(np.abs(x - y)) for x, y in zip(before, fast)) < 5e-5
    b.close()


@pytest.mark.gpu
def test_gpu_plain_loop_matches_oracle(gpu, oracle):
    work = circuits.variational_batch(12, seed=8, nmin=3, nmax=12, dmin=1, dmax=4)
    ref = pyqcs.Batch(oracle)
    b = pyqcs.Batch(gpu)
    for n, g in work:
        ref.add(n, g)
        b.add(n, g)
    want, _ = ref.run()
    got, secs = b.run_plain()
    assert secs > 0
    assert max(np.max(np.abs(x - y)) for x, y in zip(got, want)) < 1e-12
    fast, _ = b.run()
    assert max(np.max(np.abs(x - y)) for x, y in zip(fast, want)) < 2e-5
    direct, _ = b.run_direct()  # the same launches without the graph: identical bits
    assert all(np.array_equal(x, y) for x, y in zip(direct, fast))
