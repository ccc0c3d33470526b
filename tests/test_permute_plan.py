"""Permutation pass planner on the CPU (no GPU calls): tests/perm_check.cpp links the
library's planner (qsv::plan_permute_passes) and checks, for seeded random and hand-picked
permutations, that the passes compose to the requested permutation and that each pass fits
one tile of the permutation kernel.  Checked against a planted mistake: a planner that
drops the closing step of a split cycle (c_r -> c_1) builds a pass that is not a
permutation; the planner's own check then raises on the first case that splits a cycle
(the 20-qubit reversal) and the checker exits non-zero.  The GPU
tests (test_gpu_permute.py) cover the kernel's data movement."""
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
    exe = tmp_path_factory.mktemp("perm") / "perm_check"
    subprocess.run([HIPCC, "-O1", "-std=c++17", f"-I{REPO / 'include'}", f"-I{CSRC}", str(REPO / "tests" / "perm_check.cpp"),
                    f"-L{LIB.parent}", "-lqcs", f"-Wl,-rpath,{LIB.parent}", "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_permute_passes_compose(checker, seed):
    p = subprocess.run([str(checker), "200", str(seed)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert p.stdout.strip().endswith("ok")
