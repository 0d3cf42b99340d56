"""Shared fixtures.  Tests marked `gpu` need an MI355X; the rest run on any CPU."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the library's kernels")
    config.addinivalue_line("markers", "slow: full-size (30-qubit / 14-qubit DM) property checks")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    if not O.ORACLE_SO.exists():
        O.build(ref=False)
    return O.Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle import oracle as O

    if not O.reference_available():
        pytest.skip("reference csim not built here (oracle/_ref)")
    return O.Reference()


@pytest.fixture(scope="session")
def golden():
    def load(name: str):
        return np.load(GOLDEN / name, allow_pickle=False)

    return load


@pytest.fixture
def rng():
    return np.random.default_rng(1234)
