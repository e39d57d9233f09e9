import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


class Lcg:
    """The reference tests' generator: state * 6364136223846793005 +
    1442695040888963407 (wrapping), (state >> 11) / 2^53
    (crates/estimation/src/kernel.rs:132-137, search.rs:1185-1190)."""

    def __init__(self, seed):
        self.state = int(seed) & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return (self.state >> 11) / float(1 << 53)


@pytest.fixture(scope="session")
def engine():
    """A HIP context; GPU tests only."""
    from boitata_amd.engine import Engine

    e = Engine(0)
    yield e
    e.close()


def rel_err(a, b, floor):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)
