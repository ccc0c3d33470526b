"""Shared test setup: markers, library handles, helpers."""
import pathlib
import sys

import numpy as np
import pytest

REPO = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "rqs-svg_amd"))
sys.path.insert(0, str(REPO / "tests"))

import pyqcs  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libqcs.so kernels)")


@pytest.fixture(scope="session")
def oracle():
    return pyqcs.oracle_library()


@pytest.fixture(scope="session")
def gpu():
    lib = pyqcs.gpu_library()  # raises if libqcs.so is not built: no silent fallback
    return lib


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return psi / np.linalg.norm(psi)


def fidelity(a, b):
    """Normalised fidelity |<a|b>|^2 / (<a|a> <b|b>): blind to a uniform scale of either."""
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def overlap_raw(a, b):
    """Unnormalised |<a|b>|^2 against a unit-norm reference b: a norm drift of a counts."""
    return abs(np.vdot(a, b)) ** 2 / np.vdot(b, b).real
