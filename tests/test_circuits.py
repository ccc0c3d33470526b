"""Workload generators: deterministic, and the gate counts the workloads state."""
from pyqcs import circuits


def test_splitmix64_known_answer():
    # SplitMix64 reference output for seed 1234567 (first value of the published stream)
    g = circuits.SplitMix64(1234567)
    assert g.next_u64() == 6457827717110365317


def test_ghz_u3_cz_counts():
    c = circuits.ghz_u3_cz(20, 20, seed=1)
    assert len(c) == 1 + 19 + 400 + 200
    assert c == circuits.ghz_u3_cz(20, 20, seed=1)
    assert c != circuits.ghz_u3_cz(20, 20, seed=2)
    # each CZ layer is a perfect matching
    for layer in range(20):
        cz = c[20 + layer * 30 + 20: 20 + layer * 30 + 30]
        qs = [q for op in cz for q in (op[1][0], op[3][0])]
        assert sorted(qs) == list(range(20))


def test_brick_counts():
    c = circuits.brick(30, 24, seed=2, mcx=4)
    names = [op[0] for op in c]
    assert names.count("u3") == 720
    assert sum(1 for op in c if op[0] == "z") == 12 * 15 + 12 * 14
    mcx = [op for op in c if op[0] == "x"]
    assert len(mcx) == 4 and all(len(op[3]) == 3 for op in mcx)
    assert all(max(op[1] + op[3]) >= 24 for op in mcx)


def test_qft_counts():
    c = circuits.qft(33)
    assert sum(op[0] == "h" for op in c) == 33
    assert sum(op[0] == "p" for op in c) == 33 * 32 // 2
    assert sum(op[0] == "swap" for op in c) == 16
