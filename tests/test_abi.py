"""The C-ABI libraries load and export every symbol the headers declare (no GPU needed)."""
import pathlib
import re
import subprocess

import pytest

import pyqcs

REPO = pathlib.Path(__file__).resolve().parent.parent


def declared():
    c_names, cxx_names = set(), set()
    for h in ["qcs.h", "qcs_ext.h"]:
        text = (REPO / "include" / h).read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|double|qcs_simulator \*)\s*\*?\s*(qcs_\w+)\s*\(", text, re.M):
            (cxx_names if m.group(1).endswith("_cxx") else c_names).add(m.group(1))
    return c_names, cxx_names


def exported(lib_path):
    out = subprocess.run(["nm", "-D", "--defined-only", "-C", str(lib_path)], capture_output=True, text=True, check=True).stdout
    names = set()
    for line in out.splitlines():
        parts = line.split(maxsplit=2)
        if len(parts) == 3 and parts[1] in ("T", "W"):
            names.add(parts[2].split("(")[0])
    return names


def test_header_counts():
    c_names, cxx_names = declared()
    # include/qcs.h:85-148 and :18-80 of the reference: 64 C and 63 C++ entry points
    assert len([n for n in c_names if not n.startswith("qcs_ext_")]) == 64
    assert len(cxx_names) == 63


@pytest.mark.parametrize("which", ["gpu", "oracle"])
def test_exports(which):
    path = pyqcs.GPU_LIB if which == "gpu" else pyqcs.ORACLE_LIB
    names = exported(path)
    c_names, cxx_names = declared()
    missing = sorted((c_names | cxx_names) - names)
    assert not missing, missing


def test_gpu_library_loads_without_gpu():
    lib = pyqcs.gpu_library()
    for n in declared()[0]:
        assert hasattr(lib.lib, n)
