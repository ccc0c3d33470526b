"""bin/qcs, the circuit plug-in driver (reference src/qcs_main.cpp:64-201).

The same driver source is linked twice: against the oracle (oracle/_build/qcs_oracle_cli,
CPU tests) and against libqcs.so (rqs-svg_amd/bin/qcs, GPU tests). Plug-ins are the C files
under tests/plugins, built with rqs-svg_amd/compile_circuit.sh like a user's circuit.
"""
import json
import os
import pathlib
import re
import subprocess

import numpy as np
import pytest

import pyqcs
from conftest import REPO, fidelity

ORACLE_CLI = REPO / "oracle" / "_build" / "qcs_oracle_cli"
GPU_CLI = REPO / "rqs-svg_amd" / "bin" / "qcs"
COMPILE = REPO / "rqs-svg_amd" / "compile_circuit.sh"
PLUGINS = REPO / "tests" / "plugins"
LINE = re.compile(r'^\{"sample_num": (\d+), "clbits": "([01]*)", "elapsed_time": ([-+0-9.eE]+)\}$')


@pytest.fixture(scope="module")
def plugin_dir(tmp_path_factory):
    out = tmp_path_factory.mktemp("plugins")
    for src in ("ghz_measure.c", "rotations.c"):
        so = out / (src[:-2] + ".so")
        subprocess.run([str(COMPILE), str(PLUGINS / src), "-o", str(so)], check=True)
    (out / "empty.c").write_text("int nothing_here(void) { return 0; }\n")
    subprocess.run([str(COMPILE), str(out / "empty.c"), "-o", str(out / "empty.so")], check=True)
    return out


def run_cli(cli, args, n, check=True):
    env = dict(os.environ, NUM_QUBITS=str(n))
    p = subprocess.run([str(cli)] + [str(a) for a in args], env=env, capture_output=True, text=True, timeout=120)
    if check:
        assert p.returncode == 0, p.stderr
    return p


def parse(stdout):
    rows = []
    for line in stdout.strip().splitlines():
        m = LINE.match(line)
        assert m, line
        rows.append((int(m.group(1)), m.group(2), float(m.group(3))))
    return rows


def rotations_reference(n):
    with pyqcs.Simulator(pyqcs.oracle_library(), n) as s:
        for q in range(n):
            s.gate("u3", [q], [0.3 + q, 0.7 * q, 1.1 - 0.2 * q])
        for q in range(n - 1):
            s.gate("rzx", [q, q + 1], [0.4 * (q + 1)])
        s.gate("rccx", [0, 1, 2], ctrls=[3])
        return s.statevector()


def ghz_checks(rows, n, samples):
    assert [r[0] for r in rows] == list(range(samples))
    outcomes = {r[1] for r in rows}
    assert outcomes <= {"0" * n, "1" * n}
    assert len(outcomes) == 2  # both branches in `samples` shots (fails with prob 2^(1-samples))
    assert all(r[2] >= 0 for r in rows)


@pytest.mark.skipif(not ORACLE_CLI.exists(), reason="oracle CLI not built")
class TestOracleCli:
    def test_ghz_samples(self, plugin_dir):
        rows = parse(run_cli(ORACLE_CLI, ["-s", 24, "--seed", 5, plugin_dir / "ghz_measure.so"], 7).stdout)
        ghz_checks(rows, 7, 24)

    def test_seed_reproducible(self, plugin_dir):
        a = parse(run_cli(ORACLE_CLI, ["-s", 16, "--seed", 9, plugin_dir / "ghz_measure.so"], 5).stdout)
        b = parse(run_cli(ORACLE_CLI, ["--num-samples", 16, "--seed", 9, plugin_dir / "ghz_measure.so"], 5).stdout)
        assert [r[1] for r in a] == [r[1] for r in b]

    def test_saved_state_is_logical_under_any_mapping(self, plugin_dir, tmp_path):
        n = 6
        want = rotations_reference(n)
        for i, extra in enumerate([[], ["--mapping", "5,4,3,2,1,0"], ["--mapping", "2,0,5,1,3,4"],
                                   ["--mapping", "2,0,5,1,3,4", "-r"]]):
            path = tmp_path / f"sv{i}.bin"
            rows = parse(run_cli(ORACLE_CLI, extra + ["--output-statevector", path, plugin_dir / "rotations.so"], n).stdout)
            assert rows[0][1] == ""  # no clbits
            got = pyqcs.load_statevector_file(path, n)
            assert np.max(np.abs(got - want)) < 1e-12

    @pytest.mark.parametrize("args,msg", [
        (["-s", 0], "num_samples"),
        (["--mapping", "0,1,1"], "multiple times"),
        (["--mapping", "0,1"], "does not match"),
        (["--mapping", "0,1,x"], "not a qubit"),
    ])
    def test_rejects_bad_options(self, plugin_dir, args, msg):
        p = run_cli(ORACLE_CLI, args + [plugin_dir / "ghz_measure.so"], 3, check=False)
        assert p.returncode == 1 and msg in p.stderr

    def test_rejects_missing_plugin_and_symbols(self, plugin_dir, tmp_path):
        p = run_cli(ORACLE_CLI, [tmp_path / "nope.so"], 3, check=False)
        assert p.returncode == 1 and "realpath" in p.stderr
        p = run_cli(ORACLE_CLI, [plugin_dir / "empty.so"], 3, check=False)
        assert p.returncode == 1 and "circuit_init" in p.stderr
        p = run_cli(ORACLE_CLI, [], 3, check=False)
        assert p.returncode == 1


@pytest.mark.gpu
class TestGpuCli:
    def test_ghz_samples(self, plugin_dir):
        rows = parse(run_cli(GPU_CLI, ["-s", 24, "--seed", 5, plugin_dir / "ghz_measure.so"], 20).stdout)
        ghz_checks(rows, 20, 24)

    def test_same_outcomes_as_oracle_cli(self, plugin_dir):
        # one measurement per sample decides every clbit of a GHZ state; both libraries draw
        # from std::mt19937_64 seeded with the same value, one draw per undecided measurement
        if not ORACLE_CLI.exists():
            pytest.skip("oracle CLI not built")
        a = parse(run_cli(GPU_CLI, ["-s", 16, "--seed", 3, plugin_dir / "ghz_measure.so"], 12).stdout)
        b = parse(run_cli(ORACLE_CLI, ["-s", 16, "--seed", 3, plugin_dir / "ghz_measure.so"], 12).stdout)
        assert [r[1] for r in a] == [r[1] for r in b]

    @pytest.mark.parametrize("precision,tol", [(0, 1e-12), (1, 1e-5), (2, 1e-12), (3, 1e-4)])
    def test_saved_state(self, plugin_dir, tmp_path, precision, tol):
        n = 14
        want = rotations_reference(n)
        for i, extra in enumerate([[], ["--mapping", ",".join(str(q) for q in reversed(range(n)))]]):
            path = tmp_path / f"sv{precision}_{i}.bin"
            args = extra + ["--precision", precision, "--block-qubits", 4, "--output-statevector", path,
                            plugin_dir / "rotations.so"]
            run_cli(GPU_CLI, args, n)
            got = pyqcs.load_statevector_file(path, n)
            assert np.max(np.abs(got - want)) < tol
            assert fidelity(got, want) > 1 - 10 * tol
