"""Precision form on the GPU: complex64 state, gates fused into k-qubit blocks on FP32 MFMA
(precision mode 1) or BF16x3 split-precision MFMA (mode 3); complex128 blocks in FP64 (mode 2).

Bars: fidelity >= 1 - 1e-6 against the FP64 oracle (and against the GPU FP64 plain path at
sizes the oracle does not reach); max|d amp| is reported and bounded by 1e-5 (FP32 level).
Where a norm drift could hide behind the normalised fidelity (BF16x3 products), the
unnormalised |<psi128|psi64>|^2 and |sum |a|^2 - 1| are held to 1e-6 as well.
"""
import math
import sys

import numpy as np
import pytest

import pyqcs
from pyqcs import circuits
from conftest import random_state, fidelity, overlap_raw

pytestmark = pytest.mark.gpu

FID_TOL = 1e-6
AMP_TOL = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def fused_vs_oracle(gpu, oracle, n, circ, k, mapping=None, psi=None, precision=1):
    with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=k, mapping=mapping) as f:
        if psi is not None:
            f.set_statevector(psi)
        f.run(circ)
        got = f.statevector()
        stats = f.fusion_stats()
    with pyqcs.Simulator(oracle, n, mapping=mapping) as o:
        if psi is not None:
            o.set_statevector(psi)
        o.run(circ)
        want = o.statevector()
    return got, want, stats


@pytest.mark.parametrize("precision", [1, 3])
@pytest.mark.parametrize("k", [4, 5])
def test_random_all_gates(gpu, oracle, k, precision):
    for n, seed in ((6, 1), (9, 2), (13, 3)):
        circ = circuits.random_all_gates(n, 200, seed=seed, max_ctrls=3)
        say(f"fused random circuit n={n} k={k} precision={precision}")
        got, want, stats = fused_vs_oracle(gpu, oracle, n, circ, k, precision=precision)
        fid = fidelity(got, want)
        err = np.max(np.abs(got - want))
        say(f"  fidelity {fid:.12f} max|d| {err:.2e} gates {stats[0]} blocks {stats[1]}")
        assert fid >= 1 - FID_TOL and err <= AMP_TOL
        assert stats[0] == len(circ) and stats[1] <= len(circ)


@pytest.mark.parametrize("precision", [1, 3])
def test_each_gate_single(gpu, oracle, precision):
    n = 8
    for name in sorted(pyqcs.GATE_TARGETS):
        rng = np.random.default_rng(len(name))
        t = pyqcs.GATE_TARGETS[name]
        qs = [int(q) for q in rng.permutation(n)]
        targets, ctl = qs[:t], qs[t:t + 2]
        k = 1 if name == "global_phase" else pyqcs.GATE_PARAMS[name]
        params = tuple(rng.uniform(-math.pi, math.pi, size=k))
        psi = random_state(n, 5)
        got, want, _ = fused_vs_oracle(gpu, oracle, n, [(name, targets, params, ctl[:1], ctl[1:])], 5, psi=psi, precision=precision)
        assert np.max(np.abs(got - want)) <= AMP_TOL, name


def test_wide_gates(gpu, oracle):
    # gates with more qubits than the block size: controls handled as a subspace
    n = 9
    circ = [("h", (q,), (), (), ()) for q in range(n)]
    circ += [("rcccx", (0, 1, 2, 3), (), (4, 5), (6,)),
             ("x", (8,), (), (0, 1, 2, 3, 4), ()),
             ("u3", (7,), (0.3, 0.2, 0.1), (1, 2, 3, 4), (5, 6)),
             ("swap", (2, 6), (), (0, 1, 3), ())]
    for k in (4, 5):
        for precision in (1, 3):
            got, want, stats = fused_vs_oracle(gpu, oracle, n, circ, k, precision=precision)
            assert np.max(np.abs(got - want)) <= AMP_TOL


@pytest.mark.parametrize("precision", [1, 3])
def test_brick_mapping_and_fusion_counts(gpu, oracle, precision):
    n = 14
    circ = circuits.brick(n, 6, seed=4, mcx=2)
    mapping = [int(q) for q in np.random.default_rng(3).permutation(n)]
    got, want, stats = fused_vs_oracle(gpu, oracle, n, circ, 5, mapping=mapping, precision=precision)
    assert fidelity(got, want) >= 1 - FID_TOL
    # 6 layers x (14 U3 + 6-7 CZ) + 2 MCX fuse into far fewer sweeps
    assert stats[1] * 3 <= stats[0]


def test_measure_sample_in_fused_mode(gpu, oracle):
    n = 10
    circ = circuits.random_all_gates(n, 80, seed=9)
    mism = 0
    for seed in range(4):
        outs = []
        for lib, prec in ((gpu, 1), (oracle, 0)):
            with pyqcs.Simulator(lib, n, num_clbits=n, seed=seed, precision=prec) as sim:
                sim.run(circ[:40])
                a = [sim.measure_to_clbit(q, q) for q in (1, 4)]
                sim.run(circ[40:])
                b = [sim.measure(q) for q in range(n)]
                outs.append(a + b)
        mism += sum(x != y for x, y in zip(*outs))
    assert mism <= 1
    # sampling from the fused state follows |amp|^2 (FP64 prefix over the complex64 state)
    psi = random_state(8, 2)
    with pyqcs.Simulator(gpu, 8, precision=1) as sim:
        sim.set_statevector(psi)
        assert abs(sim.probability_total() - 1) < 1e-6
        out = sim.sample(40000, seed=3)
    counts = np.bincount(out.astype(np.int64), minlength=256)
    p = np.abs(psi) ** 2
    chi2 = np.sum((counts - 40000 * p) ** 2 / (40000 * p))
    assert chi2 < 256 + 6 * math.sqrt(512)


def test_window_in_fused_mode(gpu, oracle):
    n = 10
    circ = circuits.random_all_gates(n, 100, seed=12)
    res = []
    for lib, prec in ((gpu, 1), (oracle, 0)):
        with pyqcs.Simulator(lib, n, seed=2, precision=prec) as sim:
            sim.set_local_qubits(7)
            sim.run(circ)
            res.append((sim.mapping(), sim.statevector()))
    assert res[0][0] == res[1][0]
    assert fidelity(res[0][1], res[1][1]) >= 1 - FID_TOL


@pytest.mark.parametrize("precision", [1, 3])
@pytest.mark.parametrize("workload", ["brick", "qft"])
def test_unnormalised_overlap_22q(gpu, precision, workload):
    # 22 qubits, brick depth 32 (the 32-qubit benchmark family: ~7 blocks per layer) or the
    # QFT on the state of a random brick layer, stages on 2^13 tiles, against the plain
    # complex128 per-gate path: normalised and unnormalised overlap, and the norm itself
    n = 22
    prep = circuits.brick(n, 1, seed=8)
    circ = circuits.brick(n, 32, seed=9) if workload == "brick" else prep + circuits.qft(n)
    say(f"unnormalised overlap n={n} {workload} precision={precision}")
    with pyqcs.Simulator(gpu, n) as plain, pyqcs.Simulator(gpu, n, precision=precision, block_qubits=5, stage_bits=13) as fused:
        plain.run(circ)
        fused.run(circ)
        ov, md = fused.overlap(plain)
        norm64, norm128 = fused.norm(), plain.norm()
        fid = abs(ov) ** 2 / (norm64 * norm128)
        raw = abs(ov) ** 2 / norm128
        say(f"  fidelity {fid:.12f} unnormalised {raw:.12f} |norm-1| {abs(norm64 - 1):.2e} max|d| {md:.2e}")
        assert abs(norm128 - 1) <= 1e-12
        assert fid >= 1 - FID_TOL and raw >= 1 - FID_TOL and abs(norm64 - 1) <= FID_TOL
        assert md <= AMP_TOL
        psi = fused.statevector()
        assert abs(float(np.vdot(psi, psi).real) - 1) <= FID_TOL


@pytest.mark.parametrize("precision", [1, 3])
def test_overlap_with_plain_gpu_path(gpu, precision):
    n = 22
    circ = circuits.brick(n, 10, seed=7, mcx=4)
    say(f"overlap fused (precision {precision}) vs plain n=22")
    with pyqcs.Simulator(gpu, n) as plain, pyqcs.Simulator(gpu, n, precision=precision) as fused:
        plain.run(circ)
        fused.run(circ)
        ov, md = fused.overlap(plain)
        fid = abs(ov) ** 2 / (fused.norm() * plain.norm())
        say(f"  fidelity {fid:.12f} max|d| {md:.2e}")
        assert fid >= 1 - FID_TOL and md <= AMP_TOL
        ov2, md2 = plain.overlap(plain)
        assert abs(ov2 - plain.norm()) < 1e-12 and md2 == 0.0


def test_precision_mode_errors(gpu):
    sim = pyqcs.Simulator(gpu, 3, precision=1, allocate=False)
    with pytest.raises(pyqcs.QcsError):
        sim.allocate()  # fewer qubits than the smallest block
    sim.close()
    with pyqcs.Simulator(gpu, 6) as sim:
        assert sim.lib.qcs_ext_set_precision(sim.h, 1, 5) == 0  # only before allocate_memory
    for mode, k in ((4, 5), (-1, 5), (3, 6), (3, 0)):
        sim = pyqcs.Simulator(gpu, 6, allocate=False)
        assert sim.lib.qcs_ext_set_precision(sim.h, mode, k) == 0
        sim.close()


def test_fused_complex128_matches_oracle(gpu, oracle):
    # precision mode 2: complex128 state, FP64 blocks of 4 qubits; FP64 bars apply
    for n, seed in ((5, 1), (9, 2), (14, 3)):
        circ = circuits.random_all_gates(n, 150, seed=seed, max_ctrls=3)
        with pyqcs.Simulator(gpu, n, precision=2, block_qubits=4) as f:
            f.run(circ)
            got = f.statevector()
            g, b = f.fusion_stats()
        with pyqcs.Simulator(oracle, n) as o:
            o.run(circ)
            want = o.statevector()
        say(f"c128 fused n={n}: max|d| {np.max(np.abs(got - want)):.2e} blocks {b} gates {g}")
        assert np.max(np.abs(got - want)) <= 1e-13
        assert fidelity(got, want) >= 1 - 1e-12


def test_fused_complex128_qft_and_measure(gpu, oracle):
    n = 12
    circ = circuits.qft(n)
    psi = random_state(n, 4)
    res = []
    for lib, prec in ((gpu, 2), (oracle, 0)):
        with pyqcs.Simulator(lib, n, seed=3, precision=prec, block_qubits=4, num_clbits=n) as sim:
            sim.set_statevector(psi)
            sim.run(circ)
            sv = sim.statevector()
            bits = [sim.measure_to_clbit(q, q) for q in range(n)]
            res.append((sv, bits))
    assert np.max(np.abs(res[0][0] - res[1][0])) <= 1e-13
    assert res[0][1] == res[1][1]


@pytest.mark.parametrize("precision", [1, 3])
@pytest.mark.parametrize("m", [11, 12, 13, 14])
def test_stages_match_oracle(gpu, oracle, precision, m):
    # stages (several blocks per sweep on LDS-resident tiles), with gates wider than a
    # block (up to 7 qubits) falling back to single launches
    for n, seed in ((m + 4, 11), (m + 6, 12)):
        circ = circuits.random_all_gates(n, 300, seed=seed, max_ctrls=3)
        mapping = [int(q) for q in np.random.default_rng(seed).permutation(n)]
        say(f"stages n={n} m={m} precision={precision}")
        with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=5, mapping=mapping, stage_bits=m) as f:
            f.run(circ)
            got = f.statevector()
            gates, launches = f.fusion_stats()
        with pyqcs.Simulator(oracle, n, mapping=mapping) as o:
            o.run(circ)
            want = o.statevector()
        fid = fidelity(got, want)
        raw = overlap_raw(got, want)
        err = np.max(np.abs(got - want))
        say(f"  fidelity {fid:.12f} unnormalised {raw:.12f} max|d| {err:.2e} gates {gates} launches {launches}")
        assert fid >= 1 - FID_TOL and raw >= 1 - FID_TOL and err <= AMP_TOL
        assert gates == len(circ)


@pytest.mark.parametrize("precision", [1, 3])
def test_window_tile_brick_stages_vs_oracle(gpu, oracle, precision):
    # the benchmark's brick family on window tiles of 2^13 amplitudes (qubits 0..3 and a run
    # of 9 consecutive qubits), 5-qubit blocks, against the FP64 oracle at 20 qubits
    n = 20
    circ = circuits.brick(n, 12, seed=44, mcx=2)
    with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=5) as f:
        f.set_plan(13, 5, tiles=1)
        f.run(circ)
        got = f.statevector()
        gates, launches = f.fusion_stats()
    with pyqcs.Simulator(oracle, n) as o:
        o.run(circ)
        want = o.statevector()
    fid, raw, err = fidelity(got, want), overlap_raw(got, want), np.max(np.abs(got - want))
    say(f"window-tile brick n={n} precision={precision}: fidelity {fid:.12f} unnormalised {raw:.12f} max|d| {err:.2e} launches {launches}")
    assert fid >= 1 - FID_TOL and raw >= 1 - FID_TOL and err <= AMP_TOL
    assert gates == len(circ) and launches * 10 <= len(circ)


@pytest.mark.parametrize("precision", [1, 3])
def test_stages_vs_block_launches(gpu, precision):
    n = 20
    circ = circuits.brick(n, 12, seed=5, mcx=3)
    res = {}
    for m in (0, 12, 13):
        with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=5, stage_bits=m, num_clbits=2, seed=4) as f:
            f.run(circ[: len(circ) // 2])
            a = f.measure_to_clbit(7, 0)  # a flush boundary with a collapse in between
            f.run(circ[len(circ) // 2:])
            res[m] = (f.statevector(), f.fusion_stats(), a)
    base = res[0][0]
    # different block decompositions of the same circuit: FP32 and BF16x3 rounding apart
    tol = 1e-9 if precision == 1 else 1e-7
    for m in (12, 13):
        assert res[m][2] == res[0][2]
        assert fidelity(res[m][0], base) >= 1 - tol
        assert res[m][1][0] == res[0][1][0]        # same gates applied
        assert res[m][1][1] * 2 <= res[0][1][1]    # at most half the sweeps
    say(f"launches: per block {res[0][1][1]}, m=12 {res[12][1][1]}, m=13 {res[13][1][1]}")


def test_stage_bits_validation(gpu):
    with pyqcs.Simulator(gpu, 8, precision=1) as sim:
        for bad in (5, 10, 15, -1):
            assert sim.lib.qcs_ext_set_stage_bits(sim.h, bad) == 0
        for good in (0, 11, 12, 13, 14):
            sim.set_stage_bits(good)


@pytest.mark.parametrize("k", [3, 4])
def test_stages_complex128_match_oracle(gpu, oracle, k):
    # precision mode 2 with stages: FP64 blocks of k qubits on LDS tiles of 2^12; FP64 bars
    for n, seed in ((16, 21), (18, 22)):
        circ = circuits.random_all_gates(n, 250, seed=seed, max_ctrls=3) + circuits.qft(n)
        mapping = [int(q) for q in np.random.default_rng(seed).permutation(n)]
        with pyqcs.Simulator(gpu, n, precision=2, block_qubits=k, mapping=mapping) as f:
            f.run(circ)
            got = f.statevector()
            gates, launches = f.fusion_stats()
        with pyqcs.Simulator(oracle, n, mapping=mapping) as o:
            o.run(circ)
            want = o.statevector()
        say(f"c128 stages n={n}: max|d| {np.max(np.abs(got - want)):.2e} gates {gates} launches {launches}")
        assert np.max(np.abs(got - want)) <= 1e-13
        assert fidelity(got, want) >= 1 - 1e-12
        assert gates == len(circ)


@pytest.mark.parametrize("m", [11, 13])
@pytest.mark.parametrize("k", [3, 4])
def test_stages_complex128_tile_sizes(gpu, oracle, m, k):
    # complex128 stages on 2^11 and 2^13 tiles (128 and 512 threads per workgroup), grown
    # and window tiles, dense and diagonal stages (the QFT's phase ladders become tables);
    # FP64 bars against the oracle
    for n, seed, tiles in ((m + 4, 31, 0), (m + 5, 32, 1)):
        circ = circuits.random_all_gates(n, 200, seed=seed, max_ctrls=3) + circuits.qft(n)
        mapping = [int(q) for q in np.random.default_rng(seed).permutation(n)]
        psi = random_state(n, seed)
        with pyqcs.Simulator(gpu, n, precision=2, block_qubits=4, mapping=mapping) as f:
            f.set_plan(m, k, tiles=tiles)
            f.set_statevector(psi)
            f.run(circ)
            got = f.statevector()
            gates, launches = f.fusion_stats()
        with pyqcs.Simulator(oracle, n, mapping=mapping) as o:
            o.set_statevector(psi)
            o.run(circ)
            want = o.statevector()
        say(f"c128 stages m={m} k={k} n={n}: max|d| {np.max(np.abs(got - want)):.2e} gates {gates} launches {launches}")
        assert np.max(np.abs(got - want)) <= 1e-13
        assert gates == len(circ) and launches < len(circ)


@pytest.mark.parametrize("k", [3, 4])
def test_stages_complex128_diagonal(gpu, oracle, k):
    # diagonal gates only (controlled phases, rz, rzz, z, s, t, ... with positive and negative
    # controls, negative controls included): every stage is diagonal and runs as one product
    # table over the tile; then a QFT on a random state, whose controlled-phase runs give
    # all-diagonal stages between stages with dense blocks.  FP64 bars against the oracle.
    names = [nm for nm in sorted(pyqcs.GATE_TARGETS) if nm in ("z", "s", "sdg", "t", "tdg", "p", "u1", "rz", "rzz", "global_phase")]
    for n, c in ((16, None), (17, None), (16, circuits.qft(16))):
        if c is None:
            c = [g for g in circuits.random_all_gates(n, 4000, seed=40 + k + n, max_ctrls=2) if g[0] in names][:300]
            assert len(c) == 300
        psi = random_state(n, 9)
        with pyqcs.Simulator(gpu, n, precision=2, block_qubits=k) as f:
            f.set_statevector(psi)
            f.run(c)
            got = f.statevector()
            gates, launches = f.fusion_stats()
        with pyqcs.Simulator(oracle, n) as o:
            o.set_statevector(psi)
            o.run(c)
            want = o.statevector()
        say(f"c128 diagonal stages n={n} k={k}: max|d| {np.max(np.abs(got - want)):.2e} gates {gates} launches {launches}")
        assert launches < len(c) // 4  # fused into stages, not run gate by gate
        assert np.max(np.abs(got - want)) <= 1e-13
        assert gates == len(c)


@pytest.mark.parametrize("precision", [1, 3])
@pytest.mark.parametrize("m", [11, 13, 14])
def test_stages_complex64_diagonal(gpu, oracle, precision, m):
    # complex64 stages whose blocks are all diagonal run as one phase table over the tile
    # (tile bits 11, 13, 14: tables of 16-128 KiB); diagonal gates only, then a QFT, whose
    # controlled-phase runs give such stages between dense ones; random states
    n = m + 4
    names = [nm for nm in sorted(pyqcs.GATE_TARGETS) if nm in ("z", "s", "sdg", "t", "tdg", "p", "u1", "rz", "rzz", "global_phase")]
    diag = [g for g in circuits.random_all_gates(n, 4000, seed=70 + m, max_ctrls=2) if g[0] in names][:300]
    assert len(diag) == 300
    psi = random_state(n, 11)
    for c in (diag, circuits.qft(n)):
        with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=5, stage_bits=m) as f:
            f.set_statevector(psi)
            f.run(c)
            got = f.statevector()
            gates, launches = f.fusion_stats()
        with pyqcs.Simulator(oracle, n) as o:
            o.set_statevector(psi)
            o.run(c)
            want = o.statevector()
        say(f"c64 diagonal stages n={n} m={m} precision={precision}: fidelity {fidelity(got, want):.12f} max|d| {np.max(np.abs(got - want)):.2e} launches {launches}")
        assert fidelity(got, want) >= 1 - FID_TOL and np.max(np.abs(got - want)) <= AMP_TOL
        assert gates == len(c)
    assert launches < len(c) // 4


@pytest.mark.parametrize("precision", [1, 3])
def test_stage_launch_chunks(gpu, oracle, precision):
    # deep circuits on few qubits give stages with more blocks than one launch's LDS holds
    # (5 blocks of 5 qubits with 2^12 tiles): those run as several launches over the tile
    n = 16
    circ = circuits.brick(n, 40, seed=31) + circuits.qft(n)
    with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=5, stage_bits=12) as f:
        f.run(circ)
        got = f.statevector()
        gates, launches = f.fusion_stats()
    with pyqcs.Simulator(oracle, n) as o:
        o.run(circ)
        want = o.statevector()
    say(f"chunks n={n} precision={precision}: fidelity {fidelity(got, want):.12f} launches {launches}")
    assert fidelity(got, want) >= 1 - FID_TOL and np.max(np.abs(got - want)) <= AMP_TOL
    assert gates == len(circ)


@pytest.mark.parametrize("precision", [1, 3])
@pytest.mark.parametrize("k", [4, 5])
def test_stage_stream_matches_tile_kernel(gpu, oracle, precision, k):
    # Stages on 2^13 tiles of 22-33 qubit states run by default in the streaming kernel (one
    # workgroup per CU; the next tile and the next block matrix are loaded into LDS while
    # the blocks run).  With one workgroup per tile instead, every state must have the same
    # bits: brick circuits (dense blocks; 2 and 4 tiles per workgroup at 22 and 23 qubits),
    # the QFT on a prepared state (trailing diagonal tables applied as the tiles are
    # stored), and a short random circuit (stages of one or two blocks).  The 22-qubit
    # brick is also held to the FP64 oracle.
    work = [
        (22, "brick", circuits.brick(22, 6, seed=7), None),
        (23, "brick", circuits.brick(23, 4, seed=9, mcx=2), None),
        (23, "qft", circuits.qft(23), circuits.brick(23, 1, seed=8)),
        (22, "short", circuits.random_all_gates(22, 12, seed=3, max_ctrls=1), None),
    ]
    for n, name, circ, prep in work:
        outs = []
        for stream in (False, True):
            with pyqcs.Simulator(gpu, n, precision=precision, block_qubits=k, stage_bits=13) as f:
                f.set_stage_stream(stream)
                if prep:
                    f.run(prep)
                f.run(circ)
                outs.append(f.statevector())
        same = np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
        say(f"stream n={n} {name} precision={precision} k={k}: identical bits {same}")
        assert same
        if name == "brick" and n == 22:
            with pyqcs.Simulator(oracle, n) as o:
                o.run(circ)
                want = o.statevector()
            fid, raw, err = fidelity(outs[1], want), overlap_raw(outs[1], want), np.max(np.abs(outs[1] - want))
            say(f"  vs oracle: fidelity {fid:.12f} unnormalised {raw:.12f} max|d| {err:.2e}")
            assert fid >= 1 - FID_TOL and raw >= 1 - FID_TOL and err <= AMP_TOL


def test_stage_stream_validation(gpu, oracle):
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, 8) as sim:
            for bad in (2, -1):
                assert sim.lib.qcs_ext_set_stage_stream(sim.h, bad) == 0
            for good in (False, True):
                sim.set_stage_stream(good)


def test_stage_stream_33q_edge(gpu):
    # The streaming kernel keeps each lane's offsets within a tile as 32-bit indices of
    # amplitude pairs, exact up to 2^33 amplitudes: at 33 qubits (two 64 GiB complex64
    # states) a brick circuit reaching the top qubit must match the per-tile kernel exactly
    # (max |d amp| 0 over the whole state, compared on the GPU).
    n = 33
    circ = circuits.brick(n, 2, seed=17)
    sims = []
    try:
        for stream in (False, True):
            f = pyqcs.Simulator(gpu, n, seed=1, precision=3, block_qubits=5, stage_bits=13)
            sims.append(f)
            f.set_stage_stream(stream)
            f.run(circ)
            f.flush()
        ov, md = sims[1].overlap(sims[0])
        norm = sims[1].norm()
        say(f"stream 33 q: <tile|stream> {ov.real:.12f}{ov.imag:+.2e}i, max|d amp| {md:.3e}, norm {norm:.12f}")
        assert md == 0.0 and abs(ov - norm) <= 1e-9
    finally:
        for f in sims:
            f.close()


@pytest.mark.parametrize("n", [9, 13, 16])
def test_complex64_reductions_match_exact_sums(gpu, n):
    # The complex64 reductions (norm, <Z_q>) take aligned pairs of amplitudes per lane: each
    # result must equal the exact sum over the state the simulator holds (read back, FP64
    # products, math.fsum) to FP64 rounding, for a partial tile (9 qubits), one tile (13)
    # and several (16), and for <Z_0>, where the two amplitudes of a pair have opposite signs.
    import math
    psi = random_state(n, 40 + n)
    with pyqcs.Simulator(gpu, n, precision=1, block_qubits=4) as f:
        f.set_statevector(psi)
        norm = f.norm()
        ez = f.expect_z()
        state = f.statevector()
    p = (state.real.astype(np.float64) ** 2 + state.imag.astype(np.float64) ** 2)
    want = math.fsum(p)
    say(f"complex64 reductions n={n}: norm {norm!r} exact {want!r}")
    assert abs(norm - want) <= 2e-15 * want  # a tree of 2^n terms: at most ~n roundings deep
    idx = np.arange(1 << n)
    for q in range(n):
        zq = math.fsum(np.where((idx >> q) & 1, -p, p))
        assert abs(ez[q] - zq) <= 2e-15 * want, q
