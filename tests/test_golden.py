"""Oracle and GPU library against the committed golden fixtures (tests/golden/)."""
import json
import pathlib

import numpy as np
import pytest

import pyqcs

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in GOLDEN.glob("*.json"))


def load(name):
    spec = json.loads((GOLDEN / f"{name}.json").read_text())
    circ = [(op[0], tuple(op[1]), tuple(op[2]), tuple(op[3]), tuple(op[4])) for op in spec["circuit"]]
    psi = np.load(GOLDEN / f"{name}.npy", allow_pickle=False)
    return spec["num_qubits"], circ, psi


def run(lib, n, circ, **kw):
    with pyqcs.Simulator(lib, n, **kw) as sim:
        sim.run(circ)
        return sim.statevector()


def test_cases_present():
    assert len(CASES) >= 5


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_golden(oracle, name):
    n, circ, psi = load(name)
    got = run(oracle, n, circ)
    assert np.max(np.abs(got - psi)) <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_gpu_matches_golden(gpu, name):
    n, circ, psi = load(name)
    got = run(gpu, n, circ)
    assert np.max(np.abs(got - psi)) <= 1e-13
    got = run(gpu, n, circ, mapping=list(range(n))[::-1])
    assert np.max(np.abs(got - psi)) <= 1e-13
