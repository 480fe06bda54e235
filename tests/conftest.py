import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
for path in (ROOT, GOLDEN, ROOT / "oracle"):
    if str(path) not in sys.path:
        sys.path.insert(0, str(path))

import swfx  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available() -> bool:
    # The kernel driver node exists only where an AMD GPU is present; this
    # check touches no GPU runtime.
    return os.path.exists("/dev/kfd")


def pytest_collection_modifyitems(config, items):
    # GPU tests run only when selected with -m gpu on a GPU host; without a
    # device they are skipped rather than failing the CPU suite.
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU on this host")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def golden():
    cache = {}

    def load(name: str) -> dict:
        if name not in cache:
            cache[name] = swfx.read(str(GOLDEN / name))
        return cache[name]

    return load


def relative_error(actual: np.ndarray, expected: np.ndarray) -> float:
    """Max-norm relative error ||a - e||_inf / ||e||_inf."""
    scale = float(np.abs(expected).max()) if expected.size else 0.0
    if scale == 0.0:
        return float(np.abs(actual - expected).max()) if actual.size else 0.0
    return float(np.abs(actual - expected).max()) / scale


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))
