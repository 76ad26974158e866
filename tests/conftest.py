import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GROUPS = ("C2", "C3", "C4", "C6", "D2", "D3", "D4", "D6", "T", "O", "I")
GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls libzxtalori.so")


@pytest.fixture(scope="session")
def operator_fixture():
    """The reference's own operator fixture (tests/fixtures/pair-operators.npz), f64."""
    with np.load(GOLDEN / "pair-operators.npz", allow_pickle=False) as data:
        return {k: data[k].copy() for k in data.files}


def ops32(fixture, group):
    return np.ascontiguousarray(fixture[group], dtype=np.float32)
