"""Writes the golden fixtures of tests/golden/ from the textbook model (tests/textbook.py).

The reference ships no fixtures, so these vectors are produced by the independent numpy
model of the published gate matrices; each fixture stores the circuit (JSON) and the final
state vector (complex128, logical order).  Run from the repository root:

    python tests/golden/make_golden.py
"""
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent / "rqs-svg_amd"))

import textbook  # noqa: E402
from pyqcs import circuits  # noqa: E402

CASES = {
    "ghz_u3_cz_n8": (8, lambda: circuits.ghz_u3_cz(8, 3, seed=1)),
    "brick_n10_d4_mcx2": (10, lambda: circuits.brick(10, 4, seed=3, mcx=2)),
    "qft_n7": (7, lambda: [("x", (0,), (), (), ()), ("x", (3,), (), (), ())] + circuits.qft(7)),
    "random_all_gates_n6": (6, lambda: circuits.random_all_gates(6, 150, seed=2, max_ctrls=2)),
    "variational_n9_d3": (9, lambda: circuits.variational(9, 3, seed=4)),
}


def main():
    for name, (n, make) in CASES.items():
        circ = make()
        psi = textbook.run(n, circ)
        ops = [[op[0], list(op[1]), [float(x) for x in op[2]], list(op[3]), list(op[4])] for op in circ]
        (HERE / f"{name}.json").write_text(json.dumps({"num_qubits": n, "circuit": ops}))
        np.save(HERE / f"{name}.npy", psi.astype(np.complex128), allow_pickle=False)
        print(name, n, len(circ), float(np.linalg.norm(psi)))


if __name__ == "__main__":
    main()
