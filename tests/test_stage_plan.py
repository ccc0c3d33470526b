"""Stage planner invariants on the CPU (no GPU calls): tests/plan_check.cpp links the
library's planner and checks seeded random gate lists under the grown, window and automatic
tile choices -- every gate in exactly one stage, order kept inside a stage, gates inside
their stage's tile, and every pair of gates the plan reorders commuting (disjoint qubits,
or both diagonal).  Checked against a planted mistake: a planner that lets any gate pass a
deferred diagonal gate fails the last check on the first case.  It then checks the phase
tables of stages of diagonal gates (complex128 and complex64) entry by entry against the
gates' own phases computed from their definitions; a flipped RZ sign in that computation
fails the first case.  The GPU parity tests cover the resulting states."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO

CSRC = REPO / "rqs-svg_amd" / "csrc"
LIB = REPO / "rqs-svg_amd" / "lib" / "libqcs.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not LIB.exists():
        pytest.skip("libqcs.so not built")
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = tmp_path_factory.mktemp("plan") / "plan_check"
    # host code only: the planner is plain C++; hipcc for the headers' HIP types
    subprocess.run([HIPCC, "-O1", "-std=c++17", f"-I{REPO / 'include'}", f"-I{CSRC}", str(REPO / "tests" / "plan_check.cpp"),
                    f"-L{LIB.parent}", "-lqcs", f"-Wl,-rpath,{LIB.parent}", "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_stage_plans_keep_semantics(checker, seed):
    p = subprocess.run([str(checker), "40", str(seed)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert p.stdout.strip().endswith("ok")
