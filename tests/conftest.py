"""Test configuration: `gpu` marks tests that need an MI355X (run with -m gpu)."""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built _pm3d_hip.so")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def golden():
    import json
    import numpy as np
    d = Path(__file__).parent / "golden"
    arrays = dict(np.load(d / "pm3d_golden.npz"))
    meta = json.loads((d / "pm3d_golden_meta.json").read_text())
    return arrays, meta
