"""qcs_ext_sample_noalloc (sampling with the prefix levels and outcomes in the library's
file-scope scratch array) against qcs_ext_sample on the same states: identical outcomes,
two calls in a row with the second on a smaller state, plus its limits."""
import numpy as np
import pytest

import pyqcs
from pyqcs import circuits


def _check(sim, shots, seed):
    want = sim.sample(shots, seed=seed)
    total = sim.probability_total()
    got = sim.sample_noalloc(shots, seed=seed)
    assert got.dtype == np.uint64 and got.shape == (shots,)
    assert np.array_equal(got, want)
    # the scratch path leaves the simulator's own prefix untouched
    assert sim.probability_total() == total
    return got


@pytest.mark.gpu
def test_noalloc_matches_sample_twice(gpu):
    # first call: 20 qubits (complex128, random mapping); second: a smaller 13-qubit state
    # and fewer shots, so the scratch array holds a larger call's contents beforehand
    with pyqcs.Simulator(gpu, 20, seed=3, mapping=list(np.random.default_rng(5).permutation(20))) as sim:
        sim.run(circuits.ghz_u3_cz(20, 6, seed=11))
        first = _check(sim, 4096, seed=77)
    with pyqcs.Simulator(gpu, 13, seed=4) as sim:
        sim.run(circuits.brick(13, 5, seed=12))
        second = _check(sim, 1000, seed=78)
        # a repeat on the same state gives the same outcomes again
        assert np.array_equal(sim.sample_noalloc(1000, seed=78), second)
    assert first.max() < (1 << 20) and second.max() < (1 << 13)


@pytest.mark.gpu
def test_noalloc_complex64_and_limits(gpu):
    with pyqcs.Simulator(gpu, 16, seed=3, precision=3, block_qubits=5) as sim:
        sim.run(circuits.brick(16, 4, seed=13))
        # queued fused gates are refused: their flush may allocate
        with pytest.raises(pyqcs.QcsError, match="flush"):
            sim.sample_noalloc(10, seed=1)
        sim.flush()
        _check(sim, 2048, seed=79)
        with pytest.raises(pyqcs.QcsError, match="65536"):
            sim.sample_noalloc(65537, seed=1)
        assert sim.sample_noalloc(0, seed=1).size == 0
