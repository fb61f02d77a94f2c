import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def frame_errors(y: np.ndarray, r: np.ndarray, per: int):
    """Per-frame relative L2 error ||y-r||/||r||; frames with ||r|| < 1e-20 are
    judged by the absolute max instead.  Returns (max_rel, p50_rel, max_abs_zero)."""
    y = np.asarray(y, np.float64).reshape(-1, per)
    r = np.asarray(r, np.float64).reshape(-1, per)
    num = np.linalg.norm(y - r, axis=1)
    den = np.linalg.norm(r, axis=1)
    ok = den >= 1e-20
    rel = num[ok] / den[ok] if ok.any() else np.zeros(1)
    zabs = np.abs(y[~ok]).max() if (~ok).any() else 0.0
    return float(rel.max()), float(np.median(rel)), float(zabs)


@pytest.fixture(scope="session")
def rng():
    return np.random.default_rng(0x5EED)
