"""Statevector file I/O: save (src/qcs.cu:2494-2543) and the load/resume extension that
replaces the reference's unreachable loader (src/qcs.cu:2121-2159).

The file is raw little-endian complex<double> in logical order, so a state saved under one
qubit mapping loads bit-identically under any other, in either library."""
import numpy as np
import pytest

import pyqcs
from pyqcs import circuits
from conftest import random_state


def roundtrip(lib_save, lib_load, n, tmp_path, map_a, map_b, precision=0):
    path = tmp_path / "sv.bin"
    psi = random_state(n, 3)
    with pyqcs.Simulator(lib_save, n, mapping=map_a) as a:
        a.set_statevector(psi)
        a.run(circuits.brick(n, 2, seed=4))
        want = a.statevector()
        a.save(path)
    assert path.stat().st_size == 16 << n
    assert np.array_equal(pyqcs.load_statevector_file(path, n), want)
    with pyqcs.Simulator(lib_load, n, mapping=map_b, precision=precision) as b:
        b.run(circuits.brick(n, 1, seed=9))  # something to overwrite
        b.load(path)
        got = b.statevector()
    return want, got


@pytest.mark.parametrize("map_a,map_b", [(None, None), ([4, 3, 2, 1, 0, 5, 6], None), (None, [6, 0, 5, 1, 4, 2, 3])])
def test_oracle_load_roundtrip(oracle, tmp_path, map_a, map_b):
    want, got = roundtrip(oracle, oracle, 7, tmp_path, map_a, map_b)
    assert np.array_equal(got, want)


def test_oracle_load_clears_measurement_and_checks_size(oracle, tmp_path):
    n = 5
    path = tmp_path / "plus.bin"
    np.full(1 << n, 2 ** (-n / 2), dtype=np.complex128).tofile(path)
    with pyqcs.Simulator(oracle, n, num_clbits=1, seed=3) as s:
        s.measure(2)
        s.load(path)
        # a fresh measurement list: the uniform state measures q2 again with one draw, and
        # the amplitude of the other branch is still there to collapse
        assert abs(s.norm() - 1) < 1e-15
        s.measure(2)
        assert abs(s.norm() - 0.5) < 1e-15
        bad = tmp_path / "short.bin"
        np.zeros((1 << n) - 1, dtype=np.complex128).tofile(bad)
        with pytest.raises(pyqcs.QcsError):
            s.load(bad)
        np.zeros((1 << n) + 1, dtype=np.complex128).tofile(bad)
        with pytest.raises(pyqcs.QcsError):
            s.load(bad)
        with pytest.raises(pyqcs.QcsError):
            s.load(tmp_path / "missing.bin")


@pytest.mark.gpu
@pytest.mark.parametrize("map_a,map_b", [(None, None), ([4, 3, 2, 1, 0] + list(range(5, 16)), None),
                                         (None, list(range(15, -1, -1)))])
def test_gpu_load_roundtrip(gpu, oracle, tmp_path, map_a, map_b):
    want, got = roundtrip(gpu, gpu, 16, tmp_path, map_a, map_b)
    assert np.array_equal(got, want)
    # across libraries: a GPU-saved file resumes in the oracle and vice versa, bit for bit
    want2, got2 = roundtrip(gpu, oracle, 12, tmp_path, list(range(11, -1, -1)), None)
    assert np.array_equal(got2, want2)
    want3, got3 = roundtrip(oracle, gpu, 12, tmp_path, None, list(range(11, -1, -1)))
    assert np.array_equal(got3, want3)


@pytest.mark.gpu
def test_gpu_load_into_fused_state(gpu, tmp_path):
    n = 14
    want, got = roundtrip(gpu, gpu, n, tmp_path, None, None, precision=1)
    assert np.max(np.abs(got - want)) < 1e-7  # complex64 storage
    want, got = roundtrip(gpu, gpu, n, tmp_path, None, None, precision=2)
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_gpu_load_errors(gpu, tmp_path):
    n = 10
    bad = tmp_path / "short.bin"
    np.zeros((1 << n) - 1, dtype=np.complex128).tofile(bad)
    with pyqcs.Simulator(gpu, n) as s:
        with pytest.raises(pyqcs.QcsError):
            s.load(bad)
        with pytest.raises(pyqcs.QcsError):
            s.load(tmp_path / "missing.bin")
        assert abs(s.norm() - 1) < 1e-15  # state untouched by the failed loads


def _measure_then_set(lib, n, psi):
    # measure q2 (it joins the measured list), overwrite the state, then a gate on q4 and a
    # measurement of q2: with the list cleared, q2 is no implicit control of the gate and
    # is measured afresh (one engine draw) on the new state
    with pyqcs.Simulator(lib, n, seed=11) as s:
        s.gate("h", [2])
        first = s.measure(2)
        s.set_statevector(psi)
        s.gate("h", [4])
        second = s.measure(2)
        return first, second, s.statevector()


def test_oracle_set_statevector_clears_measurement(oracle):
    n = 6
    psi = np.full(1 << n, 2 ** (-n / 2), dtype=np.complex128)
    first, second, out = _measure_then_set(oracle, n, psi)
    # H on q4 of the uniform state gives the state with q4 = 0 (amplitude sqrt(2) 2^(-n/2)),
    # then q2's measurement keeps half of it
    want = np.where(((np.arange(1 << n) >> 4) & 1) == 0, 2 ** (-n / 2) * np.sqrt(2), 0.0)
    want = np.where(((np.arange(1 << n) >> 2) & 1) == second, want, 0.0)
    assert np.max(np.abs(out - want)) < 1e-15


@pytest.mark.gpu
def test_gpu_set_statevector_clears_measurement(gpu, oracle):
    n = 13
    psi = random_state(n, 21)
    g = _measure_then_set(gpu, n, psi)
    o = _measure_then_set(oracle, n, psi)
    assert g[0] == o[0] and g[1] == o[1]
    assert np.max(np.abs(g[2] - o[2])) <= 1e-13
